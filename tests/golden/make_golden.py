"""Writes the fixtures under tests/golden/.

  brute_profiles.npz   small seeded series with their exact nearest-neighbour profiles,
                       computed by brute force on explicitly z-normalised windows
                       (oracle.brute_profile: direct vector differences, no QT identity,
                       no prefix sums) -- the independent check that pins the oracle.
  series_sha256.json   SHA-256 of the float32 bytes of the workload series, so a change of
                       generator (or of numpy's PCG64 stream) is caught.
  palmad_c1.json       the oracle's PALMAD result on the 65,536-point workload (m = 128,
                       top-3), cached so GPU runs need not recompute it; the oracle itself is
                       pinned by brute_profiles.npz and the closed-form tests.

Run from the repository root:  python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from paddi_amd import synth  # noqa: E402


def brute_cases():
    cases = []
    cases.append(("rw_1200_m32", synth.rw_planted(1200, 11, 2, 40, 10.0), 32))
    cases.append(("ecg_1500_m50", synth.ecg_like(1500, 12, 1, 1), 50))
    x = synth.rw_planted(900, 13, 1, 30, 8.0).astype(np.float64)
    x[400:470] = 2.5  # a flat stretch: windows inside it are degenerate (sd == 0)
    cases.append(("flat_900_m24", x.astype(np.float32), 24))
    cases.append(("ar2_700_m16", synth.ar2_mix(700, 14), 16))
    cases.append(("rw_300_m8", synth.rw_planted(300, 15, 0, 8, 4.0), 8))
    return cases


def main():
    out = {}
    for name, x, m in brute_cases():
        nn, arg = O.brute_profile(x, m)
        out[f"{name}__x"] = x
        out[f"{name}__m"] = np.array([m])
        out[f"{name}__nn"] = nn
        out[f"{name}__arg"] = arg
    np.savez_compressed(os.path.join(HERE, "brute_profiles.npz"), **out)

    shas = {}
    for key in ("c1", "c2", "c3", "c4"):
        w = synth.WORKLOADS[key]()
        shas[key] = synth.series_sha256(w["series"][0])
    with open(os.path.join(HERE, "series_sha256.json"), "w") as f:
        json.dump(shas, f, indent=1)

    w = synth.workload_c1()
    res = O.palmad(w["series"][0], w["minL"], w["maxL"], w["step"], w["k"])
    with open(os.path.join(HERE, "palmad_c1.json"), "w") as f:
        json.dump([dict(m=r["m"], idx=[int(i) for i in r["idx"]], d2=[float(d) for d in r["d2"]],
                        attempts=r["attempts"], r2=r["r2"]) for r in res], f, indent=1)


if __name__ == "__main__":
    main()

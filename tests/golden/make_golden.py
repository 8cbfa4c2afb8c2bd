"""Write the golden fixtures of tests/golden/ from the reference's own code.

Run in a container that has the reference sources (it builds
oracle/_ref/libpmref.so from them via `make -C oracle ref`):

    python tests/golden/make_golden.py

Fixtures (inputs and outputs only):
  ref_kick_ng{NG}_np{NP}.npz  positions from the reference's init_particles
      (one rank, utils.cpp:22-36), the density contrast after the ghost fold
      (ReorderSlab), the potential after forward/apply_green(0.9)/backward
      (ffte1) and the momenta after one update_particles kick from zero,
      all computed by the reference's unmodified utils/reorder/FFTE code at
      one rank (transposes done as the one-rank copies they stand for).
  std_mt19937_64.npz  raw std::mt19937_64 outputs of libstdc++ for seeds
      0, 1, 2 and 5489 (the reference's generator, utils.cpp:23-25).
"""
import subprocess
import sys
import tempfile
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT / "oracle"))
import pyoracle  # noqa: E402

OUT = Path(__file__).resolve().parent

MT_SRC = r"""
#include <random>
#include <cstdio>
#include <cstdint>
int main() {
  const uint64_t seeds[4] = {0, 1, 2, 5489};
  for (int s = 0; s < 4; ++s) {
    std::mt19937_64 g(seeds[s]);
    for (int i = 0; i < 10000; ++i) { unsigned long long v = g(); if (i < 16 || i == 9999) std::printf("%llu\n", v); }
  }
}
"""


def main():
    subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "all", "ref"], check=True)
    for ng, nrow, keep_mesh in [(16, 16, True), (16, 12, True), (32, 16, False)]:
        npart = nrow ** 3
        pos = pyoracle.ref_init_particles(npart, ng)
        out = pyoracle.ref_step(ng, pos, a=0.9)
        data = {"ng": np.int64(ng), "a": np.float64(0.9), "posx": pos[0], "posy": pos[1], "posz": pos[2],
                "px": out["px"], "py": out["py"], "pz": out["pz"]}
        if keep_mesh:
            data["slab"] = out["slab"].reshape(ng, ng, ng)
            data["phi"] = out["phi"].reshape(ng, ng, ng)
        np.savez_compressed(OUT / f"ref_kick_ng{ng}_np{npart}.npz", **data)
    with tempfile.TemporaryDirectory() as td:
        src = Path(td) / "mt.cpp"
        exe = Path(td) / "mt"
        src.write_text(MT_SRC)
        subprocess.run(["g++", "-O2", "-o", str(exe), str(src)], check=True)
        vals = [int(v) for v in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    arr = np.array(vals, dtype=np.uint64).reshape(4, 17)
    np.savez_compressed(OUT / "std_mt19937_64.npz", seeds=np.array([0, 1, 2, 5489], dtype=np.uint64),
                        first16=arr[:, :16], out10000=arr[:, 16])
    print("fixtures written to", OUT)


if __name__ == "__main__":
    main()

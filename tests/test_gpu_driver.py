"""The one-box C++ driver (particle-mesh_amd/bin/pm_main, the reference's
main.cpp loop over the drop-in FFT / TransposeSlab / ReorderSlab /
ReorderSlabBack / TransposeSlabBack classes and utils functions) reproduces
the reference's own kick: positions never move in that loop, so after W + L
iterations the momenta are (W + L) times the golden one-kick momenta."""
import os
import subprocess
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]
EXE = ROOT / "particle-mesh_amd" / "bin" / "pm_main"


@pytest.mark.parametrize("name", ["ref_kick_ng16_np4096", "ref_kick_ng16_np1728", "ref_kick_ng32_np4096"])
def test_driver_reproduces_reference_kicks(dev, golden_dir, tmp_path, name):
    assert EXE.exists(), "build first"
    f = np.load(golden_dir / f"{name}.npz")
    ng, npart = int(f["ng"]), f["px"].size
    nrow = round(npart ** (1 / 3))
    out = tmp_path / "p.bin"
    print(f"case driver {name}", flush=True)
    r = subprocess.run([str(EXE), str(ng), str(nrow), "--warmup", "1", "--loop", "2", "--dump", str(out)],
                       capture_output=True, text=True, timeout=120, env=dict(os.environ, PM_WISDOM_DIR=str(tmp_path)))
    assert r.returncode == 0, r.stderr
    times = [float(v) for v in r.stdout.split()]
    assert len(times) == 9 and all(t >= 0 for t in times)
    p = np.fromfile(out, dtype=np.float64).reshape(3, npart)
    for k, key in enumerate(("px", "py", "pz")):
        ref = 3.0 * f[key]
        assert np.max(np.abs(p[k] - ref)) < 1e-11 * np.abs(ref).max()


def test_driver_rejects_bad_arguments():
    if not EXE.exists():
        pytest.skip("driver not built")
    r = subprocess.run([str(EXE), "7", "4"], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0

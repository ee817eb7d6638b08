"""Writes the FP64 oracle reference of the heightfield workload
(1024 x 1024 grid, 2048 x 1024 dome, 1920 x 1080, 8 bounces) at N spp
(default 4096) on every STEP-th row and column (default 8), as two independent N/2-spp
halves (sample numbers 100000.. and 100000 + N/2..), so that a reader can
estimate its noise floor. Output: c1_reference_<N>spp.npz (rows, cols,
half_a, half_b; RGB means per pixel, float64).

Run from the repository root:
    python tests/golden/make_c1_reference.py [N] [threads] [STEP]

The committed files: c1_reference_4096spp.npz (STEP 8: 135 x 240 pixels) and
c1_reference_32768spp.npz (STEP 32: 34 x 60 pixels).
"""

import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import hdcodex_amd as hx  # noqa: E402
import oracle  # noqa: E402

W, H, BOUNCES = 1920, 1080, 8


def main() -> None:
    spp = int(sys.argv[1]) if len(sys.argv) > 1 else 4096
    threads = int(sys.argv[2]) if len(sys.argv) > 2 else 8
    step = int(sys.argv[3]) if len(sys.argv) > 3 else 8
    oracle.use_fastest_build()
    half = spp // 2
    scene = hx.ProceduralScene("heightfield", 0, 1024, 2048)
    o = oracle.OracleScene(scene.view)
    cam = scene.camera(W, H)
    rows = np.arange(0, H, step, dtype=np.uint32)
    halves = []
    for base in (100000, 100000 + half):
        t = time.time()
        img = o.render(cam, W, H, base, BOUNCES, half, precision=64, rows=rows, col_step=step, threads=threads) / half
        print(f"half at {base}: {time.time() - t:.1f} s", flush=True)
        halves.append(img)
    out = Path(__file__).resolve().parent / f"c1_reference_{spp}spp.npz"
    np.savez_compressed(out, rows=rows, cols=np.arange(0, W, step), half_a=halves[0], half_b=halves[1])
    a, b = halves
    ref = 0.5 * (a + b)
    floor = float(np.mean((a - b) ** 2 / (ref * ref + 0.01))) / 4.0
    print(f"wrote {out}; estimated rel-MSE noise floor of the {spp}-spp mean {floor:.3e}")


if __name__ == "__main__":
    main()

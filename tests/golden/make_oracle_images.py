#!/usr/bin/env python3
"""Write the FP64 CPU-oracle reference images used by bench.py and the tests.

    python tests/golden/make_oracle_images.py [c1] [c2] [c4]

Each file <workload>_oracle_fp64.npz holds `image` (strided pixel subset of the
converged frame, FP64 radiance / float32 storage), `stride`, `spp`, `seed`.
The inputs are the seeded synthetic workloads of renderer2_amd.scenes; the
oracle seed (7777) is independent of the product renders' seeds.
"""
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))

import oracle as O  # noqa: E402
from renderer2_amd import scenes  # noqa: E402

PLAN = {
    # workload: (pixel stride, spp, overrides)
    "c1": (1, 4096, {}),
    "c2": (8, 4096, {}),
    "c4": (8, 4096, {"precision": "fp32"}),
}
SEED = 7777


def main(names):
    for name in names:
        stride, spp, ov = PLAN[name]
        w = scenes.make_workload(name, **ov)
        cfg = dict(w["render"], seed=SEED)
        W, H = cfg["width"], cfg["height"]
        sc = O.Scene(w["grids"], w["instances"], w["env"], "f64")
        ys, xs = np.meshgrid(np.arange(0, H, stride), np.arange(0, W, stride), indexing="ij")
        pix = (ys * W + xs).reshape(-1).astype(np.int32)
        img = np.zeros((len(pix), 3), np.float64)
        t0 = time.time()
        chunk = 256
        for s0 in range(0, spp, chunk):
            n = min(chunk, spp - s0)
            P = np.repeat(pix, n)
            S = np.tile(np.arange(s0, s0 + n, dtype=np.int32), len(pix))
            out = sc.render_paths(dict(cfg, spp=spp), P, S)
            img += out.reshape(len(pix), n, 3).sum(axis=1)
            print(f"{name}: {s0 + n}/{spp} spp, {time.time() - t0:.0f}s", flush=True)
        img = img.reshape(ys.shape + (3,)).astype(np.float32)
        np.savez_compressed(Path(__file__).parent / f"{name}_oracle_fp64.npz", image=img, stride=stride, spp=spp,
                            seed=SEED)


if __name__ == "__main__":
    main(sys.argv[1:] or ["c1", "c2"])

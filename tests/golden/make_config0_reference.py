"""Writes tests/golden/config0_ref4096.npz: the oracle's FP64-accumulated 4096-spp image
of the Cornell-box configuration (256x256, max depth 4, Russian roulette from depth 3,
seed 0), the image that whole GPU frames of that configuration are scored against by
rel-MSE. Run from the repository root (`python tests/golden/make_config0_reference.py`;
about 3 minutes on 8 cores, `--threads N` to change).

The 4096 samples are traced in chunks of 256 through the oracle's sample offset (the
per-path keys are (seed, pixel, sample, depth), so a chunked render traces exactly the
paths of one 4096-spp render); each chunk's FP64 mean is summed in FP64. Stored as
float32 (the FP64 mean rounded once), with the per-pixel second moment of the path
radiance in float16 (for the expected error of an n-spp estimate) and the configuration.
"""
import argparse
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "config0_ref4096.npz")
SPP, CHUNK = 4096, 256


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=0)
    args = ap.parse_args()
    from oracle import Oracle
    from trayracer_amd import scenes
    sc = scenes.cornell()  # 256x256, 16 spp, depth 4, RR from 3, seed 0
    r = sc["render"]
    W, H = r["width"], r["height"]
    orc = Oracle(sc)
    s1 = np.zeros((H, W, 3), np.float64)
    s2 = np.zeros((H, W, 3), np.float64)
    t0 = time.time()
    for off in range(0, SPP, CHUNK):
        out = orc.render(render=dict(spp=CHUNK), sample_offset=off, threads=args.threads)
        L = out["path_rgb"].astype(np.float64).reshape(CHUNK, H, W, 3)
        s1 += L.sum(axis=0)
        s2 += (L * L).sum(axis=0)
        print(f"samples {off + CHUNK}/{SPP} ({time.time() - t0:.0f} s)", flush=True)
    mean = s1 / SPP
    m2 = s2 / SPP
    assert np.isfinite(m2.astype(np.float16)).all()
    np.savez_compressed(OUT, mean=mean.astype(np.float32), second_moment=m2.astype(np.float16),
                        width=W, height=H, spp=SPP, max_depth=r["max_depth"], rr_depth=r["rr_depth"], seed=r["seed"])
    print("wrote", OUT)


if __name__ == "__main__":
    main()

"""Write tests/golden/golden.json and oracle_c60_views.npz.

Inputs are the reference's own data files kept here as fixtures
(DemoData/c60.uvf, c60.1dt, c60.2dt of the reference tree).  The JSON holds
the known answers stored inside c60.uvf (MD5, 1D/2D histogram digests,
max gradient, min/max block) and the TF file facts; the npz holds FP64
oracle renders of a 32^3 crop of c60 with the c60.1dt TF at three views,
a regression fixture for the oracle itself.

Run:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402

CROP = (slice(0, 32), slice(32, 64), slice(32, 64))
VIEWS = [
    dict(view_pos=(0.0, 0.0, 2.0), view_dir=(0.0, 0.0, -1.0), up_dir=(0.0, 1.0, 0.0)),
    dict(view_pos=(1.4, 0.6, 1.2), view_dir=(-1.4, -0.6, -1.2), up_dir=(0.0, 1.0, 0.0)),
    dict(view_pos=(-0.3, 1.9, 0.4), view_dir=(0.3, -1.9, -0.4), up_dir=(0.0, 0.0, 1.0)),
]


def digest(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main() -> None:
    u = O.read_uvf_v2(os.path.join(HERE, "c60.uvf"))
    ras, h1, h2, mm = u["blocks"]
    vol = ras["volume"]
    tf = np.loadtxt(os.path.join(HERE, "c60.1dt"), skiprows=1, dtype=np.float32)
    g = {
        "c60": {
            "md5": u["checksum"].hex(),
            "md5_ok": bool(u["md5_ok"]),
            "domain": ras["domain"],
            "brick_size": ras["brick_size"],
            "overlap": ras["overlap"],
            "lod_counts": ras["lod_counts"],
            "data_offset": ras["data_offset"],
            "volume_sha256": digest(vol),
            "hist1d_sha256": digest(h1["histogram"]),
            "hist1d_sum": int(h1["histogram"].sum()),
            "hist2d_sha256": digest(h2["histogram"]),
            "hist2d_shape": list(h2["histogram"].shape),
            "max_grad": float(h2["max_grad"]),
            "minmax": mm["minmax"].reshape(-1).tolist()[:2],
        },
        "c60_1dt": {"n": int(tf.shape[0]),
                    "alpha_nonzero": [int(i) for i in np.nonzero(tf[:, 3] > 0)[0][[0, -1]]]},
        "crop": [[s.start, s.stop] for s in CROP],
        "views": VIEWS,
    }
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(g, f, indent=1, sort_keys=True)
    crop = np.ascontiguousarray(vol[CROP])
    imgs, samples = [], []
    for v in VIEWS:
        p = O.make_params(crop, tf, width=96, height=96, fov=50.0, rate=1.0, brick_size=16,
                          apron=2, **v)
        img, ns, _ = O.render(p)
        imgs.append(O.to_rgba8(img))
        samples.append(ns)
    np.savez_compressed(os.path.join(HERE, "oracle_c60_views.npz"), rgba8=np.stack(imgs),
                        samples=np.stack(samples))
    print("wrote golden.json and oracle_c60_views.npz")


if __name__ == "__main__":
    main()

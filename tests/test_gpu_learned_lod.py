"""Learned LOD: the bf16 MLP on MFMA against its numpy restatement and a
PyTorch fp32 forward, and the whole pipeline (features, training data from
the renderer's own fine-vs-coarse renders, training, per-region selection)
at a small size: the selected frame stays within 1/255 RMSE of full
resolution and evaluates fewer samples."""
import math

import numpy as np
import pytest

from conftest import gpu_renderer
from imagevis3d_amd import learned_lod as LL
from imagevis3d_amd import transfer
from imagevis3d_amd.renderer import VOL_NOISE_TUBES

pytestmark = pytest.mark.gpu


def test_mlp_predictor_matches_references():
    import torch

    rng = np.random.default_rng(0)
    sizes = [(32, 64), (64, 64), (64, 64), (64, 64)]
    ws = [rng.normal(0, 1.0 / math.sqrt(i), (i, o)).astype(np.float32) for i, o in sizes]
    wo = rng.normal(0, 0.125, 64).astype(np.float32)
    wbits = np.concatenate([LL.to_bf16_bits(w).ravel() for w in ws] + [LL.to_bf16_bits(wo)])
    b = rng.normal(0, 0.1, 4 * 64 + 1).astype(np.float32)
    x = rng.normal(0, 1, (1000, 32)).astype(np.float32)
    print("case lod mlp 1000 rows", flush=True)
    pred = LL.Predictor(wbits, b)
    y = pred(x)
    ref = LL.reference_forward(wbits, b, x)
    assert np.abs(y - ref).max() <= 2e-2 * (1 + np.abs(ref).max())
    # plain fp32 forward with the bf16-rounded weights
    h = torch.from_numpy(LL.from_bf16_bits(LL.to_bf16_bits(x)))
    off = 0
    wf = torch.from_numpy(LL.from_bf16_bits(wbits))
    for k, (i, o) in enumerate(sizes):
        W = wf[off:off + i * o].reshape(i, o)
        off += i * o
        h = torch.relu(h @ W + torch.from_numpy(b[64 * k:64 * k + 64]))
    yt = (h @ wf[off:off + 64] + float(b[256])).numpy()
    assert np.abs(y - yt).max() <= 5e-2 * (1 + np.abs(yt).max())
    assert pred.kernel_ms > 0
    pred.close()


def test_mlp_ragged_row_counts():
    """Row counts around the 64-row wave and 256-row work-group tiles."""
    rng = np.random.default_rng(5)
    sizes = [(32, 64), (64, 64), (64, 64), (64, 64)]
    ws = [rng.normal(0, 1.0 / math.sqrt(i), (i, o)).astype(np.float32) for i, o in sizes]
    wo = rng.normal(0, 0.125, 64).astype(np.float32)
    wbits = np.concatenate([LL.to_bf16_bits(w).ravel() for w in ws] + [LL.to_bf16_bits(wo)])
    b = rng.normal(0, 0.1, 4 * 64 + 1).astype(np.float32)
    pred = LL.Predictor(wbits, b)
    for n in (1, 17, 63, 64, 65, 255, 256, 257, 4097):
        print(f"case lod mlp {n} rows", flush=True)
        x = rng.normal(0, 1, (n, 32)).astype(np.float32)
        y = pred(x)
        ref = LL.reference_forward(wbits, b, x)
        assert y.shape == (n,) and np.abs(y - ref).max() <= 2e-2 * (1 + np.abs(ref).max()), n
    pred.close()


def random_views(n, seed, w, h, fov=50.0):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        th, ph = rng.uniform(0, 2 * math.pi), rng.uniform(-1.0, 1.0)
        d = rng.uniform(1.2, 1.9)
        eye = (d * math.cos(ph) * math.sin(th), d * math.sin(ph), d * math.cos(ph) * math.cos(th))
        out.append(dict(view_pos=eye, view_dir=tuple(-e for e in eye), up_dir=(0.0, 1.0, 0.0),
                        fov=fov, width=w, height=h))
    return out


def test_pipeline_small():
    n, edge, n_lod = 256, 32, 4
    print("case learned LOD 256^3", flush=True)
    r = gpu_renderer(width=64, height=64)
    r.setGeneratedVolume(VOL_NOISE_TUBES, 4, (n, n, n), 8, edge, 2, n_lod)
    r.setSampleRateModifier(2.0)
    hist = LL.region_histograms(r)
    detail = LL.region_detail(r)
    assert detail.shape == (512, n_lod) and (detail[:, 0] == 0).all()
    assert (detail[:, 1:].mean(0) > 0).all()
    tfs = [transfer.tubes_tf(256, seed=s) for s in range(2)]
    X, Y = LL.training_rows(r, (n, n, n), edge, n_lod, random_views(12, 1, 160, 90), tfs, hist,
                            detail)
    assert len(X) > 2000
    wbits, b, loss = LL.train(X, Y, steps=2000)
    print(f"rows {len(X)} loss {loss:.3f}")
    pred = LL.Predictor(wbits, b)
    tf = transfer.tubes_tf(256, seed=7)
    r.set1DTrans(tf)
    mom = LL.tf_moments(hist, tf)
    dims = (n, n, n)
    safety, worst = LL.calibrate_safety(r, pred, random_views(3, 3, 320, 180), dims, edge, n_lod,
                                        hist, mom, detail=detail)
    print(f"calibrated budget factor {safety} (worst calibration rmse {worst:.3f}/255)")
    assert safety > 0
    rm = []
    for view in random_views(4, 2, 320, 180):
        r.resize((320, 180))
        LL.set_view(r, view)
        full = LL.full_resolution(r, dims, edge)
        st_full = r.frameStats()
        lev, sse, mem = LL.choose_levels(pred, view, dims, edge, n_lod, hist, mom, safety,
                                         detail=detail)
        r.setRegionLODs(lev)
        img = r.capture()
        st = r.frameStats()
        rm.append(LL.rmse255(img, full))
        print(f"rmse {rm[-1]:.3f}/255, samples {st.evaluated_samples} vs {st_full.evaluated_samples}, "
              f"levels {np.bincount(lev, minlength=n_lod)}")
        assert st.evaluated_samples <= st_full.evaluated_samples
    # held-out views: the calibrated selector stays near the 1/255 target
    assert np.mean(rm) <= 1.0 and max(rm) <= 1.5
    pred.close()
    r.cleanup()

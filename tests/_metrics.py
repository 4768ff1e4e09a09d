"""Film comparison metrics shared by the GPU parity tests."""
import numpy as np


def rel_l1_per_channel(gpu: np.ndarray, ref: np.ndarray) -> np.ndarray:
    """sum |gpu - ref| / sum |ref| for X, Y, Z separately (the vol_diff_gate
    metric of vol_photons.rs:1460-1489, split per channel)."""
    g = gpu.reshape(-1, 3).astype(np.float64)
    r = ref.reshape(-1, 3).astype(np.float64)
    den = np.abs(r).sum(0)
    return np.abs(g - r).sum(0) / np.maximum(den, 1e-300)


def diff_gate_stats(gpu_mean: np.ndarray, cpu_mean: np.ndarray) -> dict:
    """diff_gate.rs:112-144: per-pixel L1 distribution, outlier fraction, energy."""
    pp = np.abs(gpu_mean.reshape(-1, 3) - cpu_mean.reshape(-1, 3)).sum(1)
    mean_pp = pp.mean()
    srt = np.sort(pp)
    pct = lambda q: srt[min(int(len(srt) * q), len(srt) - 1)]
    ce, ge = cpu_mean.sum(), gpu_mean.sum()
    return {"p50": pct(0.5), "p99": pct(0.99), "max": srt[-1],
            "frac_above_10x_mean": float((pp > 10 * mean_pp).mean()) if mean_pp > 0 else 0.0,
            "energy_rel": abs(ce - ge) / max(ce, 1e-6)}

"""Pins the oracle's restatement of the viewer's display pass and exposure
quantile (spectral-viewer main.rs:131-174, 614-644; prism_on_plane.rs:172-176)
with closed forms and invariants (CPU only): a one-tap kernel is the plain
tonemap, a constant film is a fixed point of the normalised gather, a point
source spreads symmetrically with Gaussian weights, and the quantile is the
k-th order statistic at k = floor(fp32(n) * q)."""
import numpy as np
import pytest

M = np.array([[3.2406, -1.5372, -0.4986], [-0.9689, 1.8758, 0.0415], [0.0557, -0.2040, 1.0570]], np.float64)


def tonemap_bytes(xyz):
    """XYZ -> linear sRGB, Reinhard, clamp at 0, gamma 1/2.2, unorm8 (in
    float64: agrees with the fp32 pipeline to one byte)."""
    lin = np.asarray(xyz, np.float64) @ M.T
    t = np.maximum(lin / (1.0 + lin), 0.0)
    v = np.clip(t ** (1.0 / 2.2), 0.0, 1.0)
    return np.floor(v * 255.0 + 0.5).astype(np.int32)


def test_one_tap_kernel_is_the_plain_tonemap(oracle):
    rng = np.random.default_rng(0)
    vol = rng.uniform(0.0, 3.0, (12, 16, 3)).astype(np.float32)
    out = oracle.display_resolve(vol, None, 2, 4.0, 0.5, 0.0)  # k = 4 / 2 = 2, radius 0
    want = tonemap_bytes(vol.astype(np.float64) * 2.0 * 0.5)
    assert np.abs(out[..., :3].astype(np.int32) - want).max() <= 1
    assert (out[..., 3] == 255).all()


def test_constant_film_is_a_fixed_point_of_the_gather(oracle):
    vol = np.full((20, 24, 3), 0.25, np.float32)
    vol[..., 1] = 0.3
    a = oracle.display_resolve(vol, None, 1, 1.0, 1.0, 0.0)
    for kr in (1.0, 2.0, 3.5):
        b = oracle.display_resolve(vol, None, 1, 1.0, 1.0, kr)
        # borders included: the weights are renormalised over the pixels inside
        assert np.abs(a.astype(np.int32) - b.astype(np.int32)).max() <= 1


def test_point_source_spreads_with_gaussian_weights(oracle):
    w = h = 15
    vol = np.zeros((h, w, 3), np.float32)
    vol[7, 7] = (0.02, 0.02, 0.02)
    kr = 2.0
    out = oracle.display_resolve(vol, None, 1, 1.0, 1.0, kr).astype(np.int32)
    # symmetric under both mirrors and the transpose
    assert (out == out[::-1]).all() and (out == out[:, ::-1]).all() and (out == out.transpose(1, 0, 2)).all()
    # pixel (7+dx, 7+dy) sees the source with weight g(dx, dy) / sum of the 25 weights
    sigma = max(kr * 0.5, 0.5)
    g = lambda dx, dy: np.exp(-(dx * dx + dy * dy) / (2 * sigma * sigma))
    wsum = sum(g(dx, dy) for dx in range(-2, 3) for dy in range(-2, 3))
    for dx, dy in ((0, 0), (1, 0), (1, 1), (2, 1), (2, 2)):
        want = tonemap_bytes(np.array([0.02, 0.02, 0.02]) * g(dx, dy) / wsum)
        assert np.abs(out[7 + dy, 7 + dx, :3] - want).max() <= 1
    assert (out[7, 10, :3] == tonemap_bytes([0.0, 0.0, 0.0])).all()  # outside the 5x5 kernel


def test_surface_adds_before_exposure(oracle):
    rng = np.random.default_rng(1)
    surf = rng.uniform(0.0, 0.2, (6, 9, 3)).astype(np.float32)
    vol = np.zeros_like(surf)
    out = oracle.display_resolve(vol, surf, 1, 1.0, 3.0, 2.0)
    assert np.abs(out[..., :3].astype(np.int32) - tonemap_bytes(surf.astype(np.float64) * 3.0)).max() <= 1


@pytest.mark.parametrize("q", [0.0, 0.5, 0.995, 0.997, 1.0])
def test_quantile_is_the_order_statistic(oracle, q):
    rng = np.random.default_rng(2)
    h, w = 37, 53
    vol = rng.exponential(1.0, (h, w, 3)).astype(np.float32)
    surf = rng.uniform(0.0, 0.1, (h, w, 3)).astype(np.float32)
    total, target = 1000, 4.0e6
    got = oracle.display_quantile(vol, surf, total, target, q)
    k = np.float32(1.0) / np.float32(total) * np.float32(target)
    lum = np.sort((vol[..., 1] * k + surf[..., 1]).ravel())
    idx = min(int(np.float32(h * w) * np.float32(q)), h * w - 1)
    assert got == max(lum[idx], np.float32(1e-9))


def test_quantile_closed_form_and_floor(oracle):
    n = 1000
    vol = np.zeros((1, n, 3), np.float32)
    vol[0, :, 1] = np.arange(n)[::-1]  # Y = n-1 .. 0
    # k-th smallest of 0..999 is k; floor(1000 * 0.997) = 997 in fp32
    assert oracle.display_quantile(vol, None, 1, 1.0, 0.997) == 997.0
    assert oracle.display_quantile(vol, None, 1, 1.0, 0.0) == np.float32(1e-9)  # Y = 0 floored
    assert oracle.display_quantile(np.zeros((4, 4, 3), np.float32), None, 0, 1.0, 0.5) == np.float32(1e-9)

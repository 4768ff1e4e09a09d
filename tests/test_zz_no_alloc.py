"""spx_display_quantile_noalloc (the exposure quantile over the library's
fixed scratch array, no device allocation) against spx_display_quantile,
through the C ABI: bit-equal results on two calls in a row, the second on a
smaller film (the scratch keeps the first call's contents and must be
cleared by the call), with and without a surface layer, on a non-blocking
stream; and films above the 2^28-pixel maximum refused."""
import ctypes as C

import numpy as np
import pytest

import spectral_amd as S
from spectral_amd import _abi as A

pytestmark = pytest.mark.gpu


def _film(ctx, w, h, seed):
    rng = np.random.default_rng(seed)
    img = rng.exponential(1.0, (h, w, 3)).astype(np.float32)
    img[rng.uniform(size=(h, w)) < 0.2] = 0.0  # ties at zero, as dark film pixels have
    f = S.Film(ctx, w, h)  # filled and waited for on the null stream
    f.upload(img)          # a synchronous copy
    return f


def _both(ctx, vol, surf, total, target, q, stream):
    lib = A.lib()
    a, b = C.c_float(-1.0), C.c_float(-2.0)
    sp = surf.ptr if surf is not None else None
    rc = lib.spx_display_quantile(ctx.handle, vol.ptr, sp, vol.width, vol.height, total, target, q, C.byref(a),
                                  stream)
    assert rc == A.SPX_OK, lib.spx_last_error(ctx.handle)
    rc = lib.spx_display_quantile_noalloc(ctx.handle, vol.ptr, sp, vol.width, vol.height, total, target, q,
                                          C.byref(b), stream)
    assert rc == A.SPX_OK, lib.spx_last_error(ctx.handle)
    return a.value, b.value


def test_noalloc_quantile_equals_the_allocating_one(ctx):
    st = ctx.stream()
    try:
        big, small = _film(ctx, 1920, 1080, 1), _film(ctx, 97, 61, 2)
        surf = _film(ctx, 97, 61, 3)
        for vol, sf, q in ((big, None, 0.997), (small, None, 0.5), (small, surf, 0.995), (big, None, 1.0),
                           (small, None, 0.0)):
            a, b = _both(ctx, vol, sf, 7, 3.0, q, st)
            print(f"{vol.width}x{vol.height} q={q} surface={sf is not None}: {a} / {b}")
            assert np.float32(a).view(np.uint32) == np.float32(b).view(np.uint32)
            assert a > 0
        for f in (big, small, surf):
            f.free()
    finally:
        ctx._lib.spx_stream_destroy(ctx.handle, st)


def test_noalloc_quantile_refuses_films_above_its_maximum(ctx):
    f = S.Film(ctx, 4, 4)
    q = C.c_float(0.0)
    rc = A.lib().spx_display_quantile_noalloc(ctx.handle, f.ptr, None, 1 << 15, 1 << 14, 1, 1.0, 0.5, C.byref(q),
                                              None)
    assert rc == A.SPX_ECAP and b"2^28" in A.lib().spx_last_error(ctx.handle)
    rc = A.lib().spx_display_quantile_noalloc(ctx.handle, f.ptr, None, 4, 4, 1, 1.0, 1.5, C.byref(q), None)
    assert rc == A.SPX_EINVAL
    f.free()

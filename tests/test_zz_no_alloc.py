"""fcis_trans_rdm12s_noalloc (no device allocation, fixed per-device
scratch) against fcis_trans_rdm12s, both through the C ABI: two calls in a
row on the same GPU, the second on a smaller space, each to 1e-12 (the FP64
RDM bar); a space whose one alpha string exceeds the scratch is rejected
with FCIS_ERR_INVALID before anything runs."""

import numpy
import pytest

from conftest import has_gpu, package

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not has_gpu(), reason="needs a HIP device")]

pkg = package()


def _rdms(fn, sp, bra, ket):
    n = sp.norb
    outs = [numpy.full((n, n), numpy.nan), numpy.full((n, n), numpy.nan)] + [
        numpy.full((n,) * 4, numpy.nan) for _ in range(3)]
    st = fn(sp.handle, pkg._lib.vptr(bra), pkg._lib.vptr(ket),
            *[pkg._lib.dptr(o) for o in outs], pkg._lib.stream_ptr())
    return st, outs


def test_noalloc_rdm_matches_allocating_entry_point():
    import torch
    lib = pkg._lib.load()
    # (12o, (6,5)) needs several alpha tiles of the 128 MiB scratch; the
    # second, smaller call reuses the scratch left by the first
    for n, ne, seed in ((12, (6, 5), 1), (8, (4, 3), 2)):
        sp = pkg.get_space(n, ne)
        g = torch.Generator(device="cuda").manual_seed(seed)
        bra = torch.randn(sp.ndet, dtype=torch.float64, device="cuda", generator=g)
        ket = torch.randn(sp.ndet, dtype=torch.float64, device="cuda", generator=g)
        st0, ref = _rdms(lib.fcis_trans_rdm12s, sp, bra, ket)
        st1, got = _rdms(lib.fcis_trans_rdm12s_noalloc, sp, bra, ket)
        assert st0 == 0 and st1 == 0, pkg._lib.load().fcis_last_error()
        for a, b in zip(got, ref):
            scale = max(numpy.abs(b).max(), 1.0)
            assert numpy.isfinite(a).all()
            assert numpy.abs(a - b).max() < 1e-12 * scale
    # NULL outputs are skipped
    sp = pkg.get_space(6, (3, 3))
    x = torch.ones(sp.ndet, dtype=torch.float64, device="cuda")
    d1 = numpy.zeros((6, 6))
    assert lib.fcis_trans_rdm12s_noalloc(sp.handle, pkg._lib.vptr(x), pkg._lib.vptr(x),
                                         pkg._lib.dptr(d1), None, None, None, None,
                                         pkg._lib.stream_ptr()) == 0
    assert abs(numpy.trace(d1) - 3 * sp.ndet) < 1e-9


def test_noalloc_rdm_rejects_spaces_beyond_the_scratch():
    import torch
    lib = pkg._lib.load()
    sp = pkg.get_space(18, (9, 9))   # one alpha string: 4 * 48620 * 336 * 8 B > 128 MiB
    x = torch.zeros(16, dtype=torch.float64, device="cuda")
    d1 = numpy.zeros((18, 18))
    st = lib.fcis_trans_rdm12s_noalloc(sp.handle, pkg._lib.vptr(x), pkg._lib.vptr(x),
                                       pkg._lib.dptr(d1), None, None, None, None,
                                       pkg._lib.stream_ptr())
    assert st == -1   # FCIS_ERR_INVALID
    assert b"scratch" in lib.fcis_last_error()

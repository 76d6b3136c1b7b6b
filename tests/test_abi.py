"""CPU-side checks of the C ABI: the library loads, exports every symbol
declared in include/fcisigma.h, and its host-only entry points and argument
validation behave (no GPU needed: nothing here launches a kernel)."""

import ctypes
import math
import os
import re

import numpy

import oracle
from conftest import ROOT, package

pkg = package()


def header_symbols():
    text = open(os.path.join(ROOT, "include", "fcisigma.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fcis_[a-z0-9_]+)\s*\(", text)))


def test_every_header_symbol_is_exported_and_bound():
    lib = pkg._lib.load()
    syms = header_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(lib, s), s
        assert s in pkg._lib.SIGNATURES, s


def test_library_is_gfx950_code_object():
    # the in-tree library carries a gfx950 device code object
    data = open(pkg._lib.LIBPATH, "rb").read()
    assert b"gfx950" in data


def test_str2addr_matches_oracle():
    lib = pkg._lib.load()
    for n, k in [(8, 4), (12, 5), (16, 8)]:
        strs = oracle.make_strings(n, k)
        for a in range(0, strs.size, max(1, strs.size // 200)):
            assert lib.fcis_str2addr(n, k, int(strs[a])) == a
            assert lib.fcis_addr2str(n, k, a) == int(strs[a])


def test_absorb_h1e_matches_oracle():
    for n, ne in [(4, (2, 2)), (8, (4, 3)), (12, (6, 6))]:
        h1e, eri = pkg.synthetic.make_problem(n, ne, seed=3, with_ci=False)
        for fac in (1.0, 0.5):
            a = pkg.absorb_h1e(h1e, eri, n, ne, fac)
            b = oracle.absorb_h1e(h1e, eri, n, ne, fac)
            numpy.testing.assert_allclose(a, b, rtol=0, atol=1e-15)


def test_pair_matrix_layouts():
    n = 5
    h1e, eri = pkg.synthetic.make_problem(n, (2, 2), seed=1, with_ci=False)
    full = oracle.unpack_pair_matrix(eri, n)
    numpy.testing.assert_array_equal(pkg.pair_matrix(full, n), eri)
    npair = eri.shape[0]
    packed = eri[numpy.tril_indices(npair)]
    numpy.testing.assert_array_equal(pkg.pair_matrix(packed, n), eri)


def test_argument_validation_without_device():
    lib = pkg._lib.load()
    h = ctypes.c_void_p()
    assert lib.fcis_space_create(ctypes.byref(h), 0, 1, 1, None) == -1
    assert lib.fcis_space_create(ctypes.byref(h), 40, 1, 1, None) == -1
    assert lib.fcis_space_create(ctypes.byref(h), 6, 7, 1, None) == -1
    assert b"invalid" in lib.fcis_last_error()
    coef = numpy.zeros(100)
    assert lib.fcis_lincomb(65, 10, None, 10, pkg._lib.dptr(coef), 0.0,
                            ctypes.c_void_p(1), None) == -1
    assert lib.fcis_sigma_plan_execute(None, None, None, None) == -1
    assert lib.fcis_init_guess(10, 10, 0, 1e-3, ctypes.c_void_p(1), None) == -1
    assert lib.fcis_csf_create(ctypes.byref(h), None, 1, None) == -1
    assert lib.fcis_csf_det2csf(None, None, None, 1, None) == -1
    assert lib.fcis_csf_csf2det(None, None, None, 1, None) == -1


def test_csf_counts_match_oracle():
    import oracle.csf as oc
    for n, na, nb, sm in [(6, 3, 3, 1), (6, 3, 3, 3), (7, 4, 3, 2), (8, 4, 4, 5), (10, 5, 5, 1),
                          (16, 8, 8, 1), (5, 3, 2, 4), (6, 4, 2, 1)]:
        ref = oc.csf_shape(n, na, nb, sm)[1] if abs(na - nb) <= sm - 1 else 0
        assert pkg.csf_fci.count_all_csfs(n, na, nb, sm) == ref


def test_batch_manifest():
    man = pkg.synthetic.batch_manifest()
    assert len(man) == 64
    for n, (na, nb), seed in man:
        assert 6 <= n <= 12 and nb == n // 2 and na - nb in (0, 1)


def test_gas_string_order_is_sector_major_permutation():
    """fcis_gas_string_order: every string once, grouped by the occupation
    tuple of the subspaces in lexicographic order, colex within a group."""
    lib = pkg._lib.load()
    norb = numpy.asarray([4, 8, 4], dtype=numpy.int32)
    for k in (3, 8):
        n = math.comb(16, k)
        out = numpy.zeros(n, dtype=numpy.uint64)
        st = lib.fcis_gas_string_order(3, norb.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), k,
                                       out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
        assert st == 0
        colex = oracle.make_strings(16, k).astype(numpy.uint64)
        assert sorted(out.tolist()) == sorted(colex.tolist())
        occ = [(bin(s & 0xF).count("1"), bin((s >> 4) & 0xFF).count("1"), bin(s >> 12).count("1"))
               for s in out.tolist()]
        rank = [lib.fcis_str2addr(16, k, s) for s in out.tolist()]
        keys = list(zip(occ, rank))
        assert keys == sorted(keys)

"""The CSF oracle (oracle/csf.py) against the reference's own csfstring.c
(oracle/_ref/libcsf_ref.so, built from pyscf/lib/csf/csfstring.c) and
against spin invariants: orthonormal spin eigenvectors, <S^2> = S(S+1) for
every CSF, csf -> det -> csf identity, and the CSF diagonal equal to the
diagonal of U^T H U for the brute-force H."""

import ctypes
import itertools
import math
import os

import numpy
import pytest

import oracle
from oracle import brute
from oracle import csf as ocsf
from conftest import package

pkg = package()
REFCSF = os.path.join(os.path.dirname(oracle.__file__), "_ref", "libcsf_ref.so")
needs_ref = pytest.mark.skipif(not os.path.exists(REFCSF), reason="oracle/_ref not built")


def _ref():
    lib = ctypes.CDLL(REFCSF)
    return lib


@needs_ref
@pytest.mark.parametrize("nspin,na_nb,smult", [(2, 0, 1), (4, 0, 1), (4, 0, 3), (5, 1, 2), (6, 0, 1),
                                              (6, 2, 3), (7, 1, 4), (8, 0, 3), (8, 0, 1), (9, 3, 4)])
def test_umat_and_coupling_strings_vs_reference(nspin, na_nb, smult):
    lib = _ref()
    # choose (na, nb) with na - nb = na_nb over nspin unpaired electrons
    nb_ = (nspin - na_nb) // 2
    na_ = nspin - nb_
    g = ocsf.gentable(nspin, smult).astype(numpy.int32).ravel()
    ncsf = ocsf.count_csfs(nspin, smult)
    strs = numpy.zeros(ncsf, dtype=numpy.uint64)
    addrs = numpy.arange(ncsf, dtype=numpy.int32)
    lib.FCICSFaddrs2str(strs.ctypes.data_as(ctypes.c_void_p), addrs.ctypes.data_as(ctypes.c_void_p),
                        ctypes.c_size_t(ncsf), g.ctypes.data_as(ctypes.c_void_p),
                        ctypes.c_int(nspin), ctypes.c_int(smult - 1))
    assert numpy.array_equal(strs.astype(numpy.int64), ocsf.coupling_strings(nspin, smult))
    nup = (nspin + na_ - nb_) // 2
    dets = oracle.make_strings(nspin, nup)
    umat = numpy.ones((dets.size, ncsf))
    lib.FCICSFmakecsf(umat.ctypes.data_as(ctypes.c_void_p), dets.ctypes.data_as(ctypes.c_void_p),
                      strs.ctypes.data_as(ctypes.c_void_p), ctypes.c_int(nspin),
                      ctypes.c_size_t(dets.size), ctypes.c_size_t(ncsf),
                      ctypes.c_int(smult - 1), ctypes.c_int(na_ - nb_))
    mine = ocsf.spin_evecs(nspin, na_, nb_, smult)
    numpy.testing.assert_allclose(mine, umat, atol=1e-14)
    numpy.testing.assert_allclose(mine.T @ mine, numpy.eye(ncsf), atol=1e-12)


@needs_ref
@pytest.mark.parametrize("norb,na,nb", [(4, 2, 2), (5, 3, 2), (6, 3, 3), (6, 4, 1)])
def test_csd_order_vs_reference(norb, na, nb):
    """Our CSD mask, decoded by the reference's FCICSFddstrs2csdstrs, must
    walk (npair, pair conf, unpaired conf, spins) in ascending CSD order."""
    lib = _ref()
    mask = ocsf.csd_mask(norb, na, nb)
    nstrb = math.comb(norb, nb)
    assert numpy.array_equal(numpy.sort(mask), numpy.arange(mask.size))
    dd = numpy.array([[oracle.addr2str(norb, na, m // nstrb) for m in mask],
                      [oracle.addr2str(norb, nb, m % nstrb) for m in mask]], dtype=numpy.uint64)
    csd = numpy.empty((4, mask.size), dtype=numpy.uint64)
    lib.FCICSFddstrs2csdstrs(csd.ctypes.data_as(ctypes.c_void_p), dd.ctypes.data_as(ctypes.c_void_p),
                             ctypes.c_size_t(mask.size), ctypes.c_int(norb), ctypes.c_int(na), ctypes.c_int(nb))
    _, rows = ocsf.csd_shape(norb, na, nb)
    key = []
    for k in range(mask.size):
        npair = int(csd[0, k])
        nspin = na + nb - 2 * npair
        nup = (nspin + na - nb) // 2
        key.append((npair, oracle.str2addr(norb, npair, int(csd[1, k])),
                    oracle.str2addr(norb - npair, nspin, int(csd[2, k])),
                    oracle.str2addr(nspin, nup, int(csd[3, k]))))
    assert key == sorted(key)
    assert len(set(key)) == len(key)


@pytest.mark.parametrize("norb,na,nb,smult", [(4, 2, 2, 1), (4, 2, 2, 3), (5, 3, 2, 2), (5, 3, 2, 4),
                                             (6, 3, 3, 1), (6, 4, 2, 3)])
def test_csf_spin_invariants(norb, na, nb, smult):
    blocks, ncsf = ocsf.csf_shape(norb, na, nb, smult)
    S2 = brute.spin_square_matrix(norb, (na, nb))
    numpy.testing.assert_allclose(S2, S2.T, atol=1e-12)
    s = 0.5 * (smult - 1)
    eye = numpy.eye(ncsf)
    for k in range(ncsf):
        d = ocsf.csf2det(eye[k], norb, na, nb, smult)
        assert abs(d @ d - 1) < 1e-12
        assert abs(d @ S2 @ d - s * (s + 1)) < 1e-10
        back = ocsf.det2csf(d, norb, na, nb, smult)
        numpy.testing.assert_allclose(back, eye[k], atol=1e-12)


@pytest.mark.parametrize("norb,na,nb,smult", [(4, 2, 2, 1), (5, 3, 2, 2), (6, 3, 3, 1), (6, 3, 3, 3)])
def test_hdiag_csf_vs_brute_force(norb, na, nb, smult):
    h1e, eri = pkg.synthetic.make_problem(norb, (na, nb), seed=7, with_ci=False)
    H = brute.hamiltonian(h1e, eri, norb, (na, nb))
    _, ncsf = ocsf.csf_shape(norb, na, nb, smult)
    U = numpy.array([ocsf.csf2det(e, norb, na, nb, smult) for e in numpy.eye(ncsf)])
    ref = numpy.einsum("kd,de,ke->k", U, H, U)
    mine = ocsf.hdiag_csf(h1e, eri, norb, na, nb, smult)
    numpy.testing.assert_allclose(mine, ref, atol=1e-12)


def test_irrep_names_of_the_d2h_subgroups():
    from conftest import package
    cs = package().csf_symm
    assert cs.irrep_name2id("D2h", "B3u") == 7 and cs.irrep_name2id("D2h", "ag") == 0
    assert cs.irrep_name2id("Dooh", "B1u") == 5          # read as D2h
    assert cs.irrep_name2id("Coov", "A2") == 1           # read as C2v
    assert cs.irrep_name2id("Cs", 'A"') == 1
    for g, table in cs.IRREP_IDS.items():
        assert sorted(table.values()) == list(range(len(table)))
        for name, i in table.items():
            assert cs.irrep_id2name(g, i) == name
    # the product of two irreps is the XOR of their ids: B1 x B2 = A2 in C2v
    assert cs.irrep_name2id("C2v", "B1") ^ cs.irrep_name2id("C2v", "B2") == cs.irrep_name2id("C2v", "A2")
    with pytest.raises(ValueError):
        cs.irrep_name2id("C2v", "Eg")
    with pytest.raises(ValueError):
        cs.irrep_name2id("Oh", "A1g")

"""ORACLE -- test infrastructure only.

A CPU restatement of the direct-CI sigma path (pyscf.fci.direct_spin1, the
function pyscf-forge's CSF and GAS solvers call: pyscf/csf_fci/csf.py:449,
:526; pyscf/mcscf/fci_gas.py:137), written in plain C (``fci_oracle.c``) and
bound here with ctypes, plus a from-scratch Slater-Condon brute force
(``brute.py``) and the reference's own libfci_gas built from its sources
(``_ref/``, see ``Makefile``).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg may import this package, and only to check or to time the CPU path.  The
product package never imports it.

Pinning status of every function is stated in ``fci_oracle.c``'s header; the
checks live in ``tests/test_oracle.py``.
"""

import ctypes
import os

import numpy

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBORC = os.path.join(_HERE, "liborc.so")
LIBREF = os.path.join(_HERE, "_ref", "libfci_gas_ref.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)
_up = ctypes.POINTER(ctypes.c_uint64)

_lib = None
_ref = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBORC):
            raise RuntimeError("oracle library missing: run `make -C oracle`")
        lib = ctypes.CDLL(LIBORC)
        lib.orc_num_strings.restype = ctypes.c_int64
        lib.orc_num_strings.argtypes = [ctypes.c_int, ctypes.c_int]
        lib.orc_str2addr.restype = ctypes.c_int64
        lib.orc_str2addr.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint64]
        lib.orc_addr2str.restype = ctypes.c_uint64
        lib.orc_addr2str.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int64]
        lib.orc_make_strings.argtypes = [ctypes.c_int, ctypes.c_int, _up]
        lib.orc_nlink.argtypes = [ctypes.c_int, ctypes.c_int]
        lib.orc_gen_linkstr_index.argtypes = [ctypes.c_int, ctypes.c_int,
                                              ctypes.c_int, _ip]
        lib.orc_absorb_h1e.argtypes = [ctypes.c_int, ctypes.c_int,
                                       ctypes.c_double, _dp, _dp, _dp]
        lib.orc_contract_2e.argtypes = [ctypes.c_int] * 3 + [_dp, _dp, _dp]
        lib.orc_contract_2e_sample.argtypes = [ctypes.c_int] * 3 + [_dp, _dp, _dp,
                                                                     ctypes.c_int64]
        lib.orc_make_hdiag.argtypes = [ctypes.c_int] * 3 + [_dp, _dp, _dp]
        lib.orc_trans_rdm12s.argtypes = [ctypes.c_int] * 3 + [_dp] * 7
        lib.orc_sigma_rows.argtypes = [ctypes.c_int] * 3 + [_dp] * 3 + [
            ctypes.c_int64, ctypes.POINTER(ctypes.c_int64), _dp]
        lib.orc_det_sigma_rows.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_int64, _up, _up, _dp,
                                           ctypes.c_int64, ctypes.POINTER(ctypes.c_int64), _dp]
        lib.orc_det_hamiltonian.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_int64, _up, _up, _dp]
        _lib = lib
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _c64(a, shape=None):
    a = numpy.ascontiguousarray(a, dtype=numpy.float64)
    if shape is not None:
        a = a.reshape(shape)
    return a


def _unpack_nelec(nelec):
    if isinstance(nelec, (int, numpy.integer)):
        nb = int(nelec) // 2
        return int(nelec) - nb, nb
    return int(nelec[0]), int(nelec[1])


def num_strings(n, k):
    return int(_load().orc_num_strings(n, k))


def str2addr(n, k, s):
    return int(_load().orc_str2addr(n, k, int(s)))


def addr2str(n, k, a):
    return int(_load().orc_addr2str(n, k, int(a)))


def make_strings(n, k):
    out = numpy.empty(num_strings(n, k), dtype=numpy.uint64)
    _load().orc_make_strings(n, k, out.ctypes.data_as(_up))
    return out


def gen_linkstr_index(n, k, tril=False):
    """int32 [nstr, nlink, 4] in pyscf's cistring.gen_linkstr_index format."""
    lib = _load()
    nlink = lib.orc_nlink(n, k)
    out = numpy.empty((num_strings(n, k), nlink, 4), dtype=numpy.int32)
    lib.orc_gen_linkstr_index(n, k, 1 if tril else 0, out.ctypes.data_as(_ip))
    return out


def pair_matrix(eri, norb):
    """(npair, npair) view of an eri given as a pair matrix or a 4-index tensor."""
    eri = numpy.asarray(eri, dtype=numpy.float64)
    npair = norb * (norb + 1) // 2
    if eri.shape == (npair, npair):
        return numpy.ascontiguousarray(eri)
    if eri.shape == (norb,) * 4:
        idx = numpy.tril_indices(norb)
        return numpy.ascontiguousarray(eri[idx[0], idx[1]][:, idx[0], idx[1]])
    raise ValueError("unsupported eri shape %s" % (eri.shape,))


def unpack_pair_matrix(eri_pair, norb):
    """4-index (norb,)*4 tensor from a 4-fold pair matrix."""
    p, q = numpy.indices((norb, norb))
    pq = numpy.maximum(p, q) * (numpy.maximum(p, q) + 1) // 2 + numpy.minimum(p, q)
    return eri_pair[pq[:, :, None, None], pq[None, None, :, :]]


def absorb_h1e(h1e, eri, norb, nelec, fac=1):
    ne = sum(_unpack_nelec(nelec))
    h1e = _c64(h1e, (norb, norb))
    eri = pair_matrix(eri, norb)
    out = numpy.empty_like(eri)
    _load().orc_absorb_h1e(norb, ne, float(fac), _d(h1e), _d(eri), _d(out))
    return out


def contract_2e(h2e, fcivec, norb, nelec):
    na, nb = _unpack_nelec(nelec)
    h2e = pair_matrix(h2e, norb)
    c = _c64(fcivec)
    shape = c.shape
    c = c.reshape(-1)
    if c.size != num_strings(norb, na) * num_strings(norb, nb):
        raise ValueError("CI vector size mismatch")
    out = numpy.empty_like(c)
    _load().orc_contract_2e(norb, na, nb, _d(h2e), _d(c), _d(out))
    return out.reshape(shape)


def contract_2e_sample(h2e, fcivec, norb, nelec, ncols):
    """Time-sampling variant of contract_2e: only the first ``ncols`` beta
    columns' blocks are processed (a partial sigma, for CPU timing)."""
    na, nb = _unpack_nelec(nelec)
    h2e = pair_matrix(h2e, norb)
    c = _c64(fcivec).reshape(-1)
    out = numpy.empty_like(c)
    _load().orc_contract_2e_sample(norb, na, nb, _d(h2e), _d(c), _d(out), int(ncols))
    return out


def sigma_rows(h1e, eri, ci, norb, nelec, dets):
    """(H c)[d] for the determinant indices ``dets`` (alpha * nstrb + beta)
    by the Slater-Condon rules on the un-absorbed integrals (fci_oracle.c
    orc_sigma_rows): spot checks of sigma where no full CPU sigma fits."""
    na, nb = _unpack_nelec(nelec)
    h1e = _c64(h1e, (norb, norb))
    eri = pair_matrix(eri, norb)
    c = _c64(ci).reshape(-1)
    if c.size != num_strings(norb, na) * num_strings(norb, nb):
        raise ValueError("CI vector size mismatch")
    dets = numpy.ascontiguousarray(dets, dtype=numpy.int64).reshape(-1)
    if dets.size and (dets.min() < 0 or dets.max() >= c.size):
        raise ValueError("determinant index out of range")
    out = numpy.empty(dets.size)
    _load().orc_sigma_rows(norb, na, nb, _d(h1e), _d(eri), _d(c), dets.size,
                           dets.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), _d(out))
    return out


def det_sigma_rows(h1e, eri, ci, norb, stra, strb, rows):
    """(H c)[rows] for the determinants (stra[j], strb[j]) of an explicit
    list (up to 64 orbitals), H restricted to the list, by the Slater-Condon
    rules on the un-absorbed integrals (fci_oracle.c orc_det_sigma_rows):
    for a GAS determinant list this is the restricted sigma P H P c."""
    h1e = _c64(h1e, (norb, norb))
    eri = pair_matrix(eri, norb)
    c = _c64(ci).reshape(-1)
    sa = numpy.ascontiguousarray(stra, dtype=numpy.uint64).reshape(-1)
    sb = numpy.ascontiguousarray(strb, dtype=numpy.uint64).reshape(-1)
    if not (c.size == sa.size == sb.size):
        raise ValueError("determinant list and CI vector differ in size")
    rows = numpy.ascontiguousarray(rows, dtype=numpy.int64).reshape(-1)
    out = numpy.empty(rows.size)
    st = _load().orc_det_sigma_rows(int(norb), _d(h1e), _d(eri), c.size,
                                    sa.ctypes.data_as(_up), sb.ctypes.data_as(_up), _d(c),
                                    rows.size, rows.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                    _d(out))
    if st != 0:
        raise ValueError("orc_det_sigma_rows failed (%d)" % st)
    return out


def det_hamiltonian(h1e, eri, norb, stra, strb):
    """Dense H over an explicit determinant list (orc_det_hamiltonian; the
    rules of det_sigma_rows), for exact restricted spectra of small lists."""
    h1e = _c64(h1e, (norb, norb))
    eri = pair_matrix(eri, norb)
    sa = numpy.ascontiguousarray(stra, dtype=numpy.uint64).reshape(-1)
    sb = numpy.ascontiguousarray(strb, dtype=numpy.uint64).reshape(-1)
    H = numpy.empty((sa.size, sa.size))
    st = _load().orc_det_hamiltonian(int(norb), _d(h1e), _d(eri), sa.size,
                                     sa.ctypes.data_as(_up), sb.ctypes.data_as(_up), _d(H))
    if st != 0:
        raise ValueError("orc_det_hamiltonian failed (%d)" % st)
    return H


_MASK64 = (1 << 64) - 1
_GOLDEN64 = 0x9E3779B97F4A7C15


def splitmix64_mix(z):
    """The SplitMix64 output function (Vigna's splitmix64.c; Stafford's
    variant 13), on Python ints."""
    z &= _MASK64
    z ^= z >> 30
    z = (z * 0xBF58476D1CE4E5B9) & _MASK64
    z ^= z >> 27
    z = (z * 0x94D049BB133111EB) & _MASK64
    return z ^ (z >> 31)


def init_guess(hdiag, nroots, noise_norm=1e-3, addrs=None):
    """Initial vectors of the reference's GAS/CAS solver: a unit vector at
    each of the nroots lowest diagonal entries, in numpy.argpartition order
    (_lowest_hdiag_addresses, pyscf/mcscf/fci_gas.py:46-58; the whole range
    when nroots >= size), each plus a global sign noise of total norm
    noise_norm, sign +1 where bit 0 of mix(i + (root + 1) * golden) is set
    (_add_distributed_noise :61-88, get_init_guess :881-902).  ``addrs``
    replaces the selected addresses (to check the noise alone)."""
    hdiag = numpy.asarray(hdiag, dtype=numpy.float64).reshape(-1)
    n = hdiag.size
    if addrs is not None:
        addrs = numpy.asarray(addrs, dtype=numpy.int64)
    elif n <= nroots:
        addrs = numpy.arange(n)
    else:
        addrs = numpy.argpartition(hdiag, nroots - 1)[:nroots]
    scale = noise_norm / numpy.sqrt(float(n))
    out = []
    idx = numpy.arange(n, dtype=numpy.uint64)
    for root, a in enumerate(addrs):
        z = idx + numpy.uint64(((root + 1) * _GOLDEN64) & _MASK64)
        z ^= z >> numpy.uint64(30)
        z *= numpy.uint64(0xBF58476D1CE4E5B9)
        z ^= z >> numpy.uint64(27)
        z *= numpy.uint64(0x94D049BB133111EB)
        z ^= z >> numpy.uint64(31)
        g = numpy.zeros(n)
        g[int(a)] = 1.0
        g += scale * numpy.where(z & numpy.uint64(1), 1.0, -1.0)
        out.append(g)
    return addrs, out


def precond(dx, hdiag, e, level_shift):
    """The reference's diagonal preconditioner (pyscf/mcscf/fci_gas.py:
    1337-1342): dx / (hdiag - e), every denominator of magnitude below
    level_shift replaced by +level_shift."""
    d = numpy.asarray(hdiag, dtype=numpy.float64) - e
    d = numpy.where(numpy.abs(d) < level_shift, level_shift, d)
    return numpy.asarray(dx, dtype=numpy.float64) / d


def make_hdiag(h1e, eri, norb, nelec):
    na, nb = _unpack_nelec(nelec)
    h1e = _c64(h1e, (norb, norb))
    eri = pair_matrix(eri, norb)
    out = numpy.empty(num_strings(norb, na) * num_strings(norb, nb))
    _load().orc_make_hdiag(norb, na, nb, _d(h1e), _d(eri), _d(out))
    return out


def trans_rdm12s(bra, ket, norb, nelec):
    """((dm1a, dm1b), (dm2aa, dm2ab, dm2bb)) with pyscf conventions."""
    na, nb = _unpack_nelec(nelec)
    bra = _c64(bra).reshape(-1)
    ket = _c64(ket).reshape(-1)
    n2 = (norb, norb)
    n4 = (norb,) * 4
    dm1a, dm1b = numpy.empty(n2), numpy.empty(n2)
    dm2aa, dm2ab, dm2bb = numpy.empty(n4), numpy.empty(n4), numpy.empty(n4)
    _load().orc_trans_rdm12s(norb, na, nb, _d(bra), _d(ket), _d(dm1a),
                             _d(dm1b), _d(dm2aa), _d(dm2ab), _d(dm2bb))
    return (dm1a, dm1b), (dm2aa, dm2ab, dm2bb)


def make_rdm12s(ci, norb, nelec):
    return trans_rdm12s(ci, ci, norb, nelec)


def make_rdm12(ci, norb, nelec):
    """Spin-traced (dm1, dm2); dm2 = aa + bb + ab + ab^T(2,3,0,1)
    (assembly as in pyscf/mcscf/fci_gas.py:724-733)."""
    (dm1a, dm1b), (dm2aa, dm2ab, dm2bb) = make_rdm12s(ci, norb, nelec)
    return dm1a + dm1b, dm2aa + dm2bb + dm2ab + dm2ab.transpose(2, 3, 0, 1)


def energy(h1e, eri, ci, norb, nelec):
    h2e = absorb_h1e(h1e, eri, norb, nelec, 0.5)
    ci = _c64(ci).reshape(-1)
    return float(ci @ contract_2e(h2e, ci, norb, nelec).reshape(-1))


def lowest_eigenvalues(h1e, eri, norb, nelec, nroots=1, tol=1e-12):
    """Lowest eigenpairs of H by ARPACK Lanczos over the oracle sigma
    (an eigensolver independent of any Davidson implementation)."""
    import scipy.sparse.linalg as sla
    na, nb = _unpack_nelec(nelec)
    ndet = num_strings(norb, na) * num_strings(norb, nb)
    h2e = absorb_h1e(h1e, eri, norb, nelec, 0.5)
    if ndet <= 600:
        from oracle import brute
        H = brute.hamiltonian(h1e, eri, norb, (na, nb))
        w, v = numpy.linalg.eigh(H)
        return w[:nroots], v[:, :nroots].T
    op = sla.LinearOperator((ndet, ndet), dtype=numpy.float64,
                            matvec=lambda x: contract_2e(h2e, x, norb, (na, nb)))
    hd = make_hdiag(h1e, eri, norb, (na, nb))
    v0 = numpy.zeros(ndet)
    v0[numpy.argmin(hd)] = 1.0
    v0 += 1e-3 * numpy.random.default_rng(7).normal(size=ndet)
    w, v = sla.eigsh(op, k=nroots, which="SA", tol=tol, v0=v0)
    order = numpy.argsort(w)
    return w[order], v[:, order].T


# --------------------------------------------------------------------------
# The reference's own libfci_gas (compiled from /root/reference sources by
# oracle/Makefile into oracle/_ref/).  Absent if it was never built.

class RefLib:
    """ctypes view of oracle/_ref/libfci_gas_ref.so (ref_harness.c)."""

    def __init__(self, path=LIBREF):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        lib = ctypes.CDLL(path)
        lib.refh_plan_create.restype = ctypes.c_void_p
        lib.refh_plan_create.argtypes = [ctypes.c_int] * 3 + [_dp]
        lib.refh_plan_execute.argtypes = [ctypes.c_void_p, _dp, _dp]
        lib.refh_plan_free.argtypes = [ctypes.c_void_p]
        lib.refh_contract_2e.argtypes = [ctypes.c_int] * 3 + [_dp, _dp, _dp]
        lib.refh_make_hdiag.argtypes = [ctypes.c_int] * 3 + [_dp, _dp, _dp]
        lib.refh_trans_rdm12s.argtypes = [ctypes.c_int] * 3 + [_dp] * 7
        lib.refh_link_table.argtypes = [ctypes.c_int] * 4 + [_ip]
        lib.refh_str2addr.restype = ctypes.c_long
        lib.refh_str2addr.argtypes = [ctypes.c_int] * 4 + [ctypes.c_ulonglong]
        self.lib = lib

    def contract_2e(self, h2e, ci, norb, nelec):
        na, nb = _unpack_nelec(nelec)
        h2e = pair_matrix(h2e, norb)
        c = _c64(ci).reshape(-1)
        out = numpy.zeros_like(c)
        st = self.lib.refh_contract_2e(norb, na, nb, _d(h2e), _d(c), _d(out))
        if st != 0:
            raise RuntimeError("libfci_gas contract failed with status %d" % st)
        return out

    def make_hdiag(self, h1e, eri, norb, nelec):
        na, nb = _unpack_nelec(nelec)
        h1e = _c64(h1e, (norb, norb))
        eri = pair_matrix(eri, norb)
        out = numpy.empty(num_strings(norb, na) * num_strings(norb, nb))
        st = self.lib.refh_make_hdiag(norb, na, nb, _d(h1e), _d(eri), _d(out))
        if st != 0:
            raise RuntimeError("libfci_gas hdiag failed with status %d" % st)
        return out

    def trans_rdm12s(self, bra, ket, norb, nelec):
        na, nb = _unpack_nelec(nelec)
        bra = _c64(bra).reshape(-1)
        ket = _c64(ket).reshape(-1)
        n2, n4 = (norb, norb), (norb,) * 4
        d = [numpy.zeros(n2), numpy.zeros(n2), numpy.zeros(n4),
             numpy.zeros(n4), numpy.zeros(n4)]
        st = self.lib.refh_trans_rdm12s(norb, na, nb, _d(bra), _d(ket),
                                        *[_d(x) for x in d])
        if st != 0:
            raise RuntimeError("libfci_gas rdm failed with status %d" % st)
        return (d[0], d[1]), (d[2], d[3], d[4])

    def link_table(self, norb, nelec, spin):
        na, nb = _unpack_nelec(nelec)
        k = na if spin == 0 else nb
        nstr = num_strings(norb, k)
        nlink = k + k * (norb - k)
        out = numpy.zeros((nstr, nlink, 4), dtype=numpy.int32)
        got = self.lib.refh_link_table(norb, na, nb, spin, out.ctypes.data_as(_ip))
        if got != nlink:
            raise RuntimeError("unexpected link count %d" % got)
        return out

    def str2addr(self, norb, nelec, spin, s):
        na, nb = _unpack_nelec(nelec)
        return int(self.lib.refh_str2addr(norb, na, nb, spin, int(s)))

    def plan(self, h2e, norb, nelec):
        """Reusable sigma plan -> callable(ci) (used for CPU timing)."""
        na, nb = _unpack_nelec(nelec)
        h2e = pair_matrix(h2e, norb)
        handle = self.lib.refh_plan_create(norb, na, nb, _d(h2e))
        if not handle:
            raise RuntimeError("libfci_gas plan creation failed")
        lib = self.lib

        class _Plan:
            def __call__(self, ci, out):
                st = lib.refh_plan_execute(handle, _d(ci), _d(out))
                if st != 0:
                    raise RuntimeError("plan execute failed %d" % st)
                return out

            def close(self):
                lib.refh_plan_free(handle)

        return _Plan()

    def gas(self, gas_orbs, nelec, blocks):
        """The reference's own restricted space (gas_space_from_blocks) for
        canonical blocks D; see RefGas."""
        return RefGas(self.lib, gas_orbs, nelec, blocks)


class RefGas:
    """ctypes view of one libfci_gas space built from blocks (ref_harness.c
    refh_gas_*): layout, sector strings, sigma, hdiag and RDMs."""

    def __init__(self, lib, gas_orbs, nelec, blocks):
        _l = ctypes.c_long
        _lp = ctypes.POINTER(ctypes.c_long)
        lib.refh_gas_create.restype = ctypes.c_void_p
        lib.refh_gas_create.argtypes = [ctypes.c_int, _ip, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_int, _ip]
        lib.refh_gas_free.argtypes = [ctypes.c_void_p]
        lib.refh_gas_info.argtypes = [ctypes.c_void_p, _lp, _ip, _ip]
        lib.refh_gas_blocks.argtypes = [ctypes.c_void_p, _lp, _ip, _ip, _lp, _lp]
        lib.refh_gas_sectors.argtypes = [ctypes.c_void_p, _ip, _lp]
        lib.refh_gas_sector_strings.argtypes = [ctypes.c_void_p, ctypes.c_int, _up]
        lib.refh_gas_contract.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
        lib.refh_gas_hdiag.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
        lib.refh_gas_rdm12s.argtypes = [ctypes.c_void_p] + [_dp] * 7
        self.lib = lib
        self.norb = tuple(int(x) for x in gas_orbs)
        self.ngas = len(self.norb)
        self.norb_total = sum(self.norb)
        self.nelec = _unpack_nelec(nelec)
        norb_a = numpy.ascontiguousarray(self.norb, dtype=numpy.int32)
        occ = numpy.ascontiguousarray(blocks, dtype=numpy.int32).reshape(-1)
        nblock = occ.size // (2 * self.ngas)
        self.handle = lib.refh_gas_create(self.ngas, norb_a.ctypes.data_as(_ip),
                                          self.nelec[0], self.nelec[1], nblock,
                                          occ.ctypes.data_as(_ip))
        if not self.handle:
            raise RuntimeError("gas_space_from_blocks failed")
        nd, nbk, ns = _l(), ctypes.c_int32(), ctypes.c_int32()
        lib.refh_gas_info(self.handle, ctypes.byref(nd), ctypes.byref(nbk), ctypes.byref(ns))
        self.ndet, self.nblock, self.nsector = nd.value, nbk.value, ns.value
        off = numpy.zeros(self.nblock, dtype=numpy.int64)
        sa = numpy.zeros(self.nblock, dtype=numpy.int32)
        sb = numpy.zeros(self.nblock, dtype=numpy.int32)
        na = numpy.zeros(self.nblock, dtype=numpy.int64)
        nb = numpy.zeros(self.nblock, dtype=numpy.int64)
        lib.refh_gas_blocks(self.handle, off.ctypes.data_as(_lp), sa.ctypes.data_as(_ip),
                            sb.ctypes.data_as(_ip), na.ctypes.data_as(_lp),
                            nb.ctypes.data_as(_lp))
        self.block_desc = [dict(offset=int(a), sa=int(b), sb=int(c), na=int(d), nb=int(e))
                           for a, b, c, d, e in zip(off, sa, sb, na, nb)]
        socc = numpy.zeros((self.nsector, self.ngas), dtype=numpy.int32)
        snstr = numpy.zeros(self.nsector, dtype=numpy.int64)
        lib.refh_gas_sectors(self.handle, socc.ctypes.data_as(_ip), snstr.ctypes.data_as(_lp))
        self.sectors = [tuple(int(v) for v in r) for r in socc]
        self.sector_nstr = [int(x) for x in snstr]

    def sector_strings(self, sid):
        out = numpy.zeros(self.sector_nstr[sid], dtype=numpy.uint64)
        self.lib.refh_gas_sector_strings(self.handle, int(sid), out.ctypes.data_as(_up))
        return out.astype(numpy.int64)

    def contract_2e(self, h2e, ci):
        h2e = pair_matrix(h2e, self.norb_total)
        c = _c64(ci).reshape(-1)
        out = numpy.zeros_like(c)
        st = self.lib.refh_gas_contract(self.handle, _d(h2e), _d(c), _d(out))
        if st != 0:
            raise RuntimeError("libfci_gas contract failed with status %d" % st)
        return out

    def plan(self, h2e):
        """Reusable restricted sigma plan -> callable(ci, out) with close()
        (used for CPU timing)."""
        lib = self.lib
        lib.refh_gas_plan_create.restype = ctypes.c_void_p
        lib.refh_gas_plan_create.argtypes = [ctypes.c_void_p, _dp]
        lib.refh_gas_plan_execute.argtypes = [ctypes.c_void_p, _dp, _dp]
        lib.refh_gas_plan_free.argtypes = [ctypes.c_void_p]
        h2e = pair_matrix(h2e, self.norb_total)
        handle = lib.refh_gas_plan_create(self.handle, _d(h2e))
        if not handle:
            raise RuntimeError("libfci_gas restricted plan creation failed")

        class _Plan:
            def __call__(self, ci, out):
                st = lib.refh_gas_plan_execute(handle, _d(ci), _d(out))
                if st != 0:
                    raise RuntimeError("plan execute failed %d" % st)
                return out

            def close(self):
                lib.refh_gas_plan_free(handle)

        return _Plan()

    def make_hdiag(self, h1e, eri):
        h1e = _c64(h1e, (self.norb_total, self.norb_total))
        eri = pair_matrix(eri, self.norb_total)
        out = numpy.empty(self.ndet)
        st = self.lib.refh_gas_hdiag(self.handle, _d(h1e), _d(eri), _d(out))
        if st != 0:
            raise RuntimeError("libfci_gas hdiag failed with status %d" % st)
        return out

    def trans_rdm12s(self, bra, ket):
        n = self.norb_total
        bra = _c64(bra).reshape(-1)
        ket = _c64(ket).reshape(-1)
        d = [numpy.zeros((n, n)), numpy.zeros((n, n))] + [numpy.zeros((n,) * 4)
                                                          for _ in range(3)]
        st = self.lib.refh_gas_rdm12s(self.handle, _d(bra), _d(ket), *[_d(x) for x in d])
        if st != 0:
            raise RuntimeError("libfci_gas rdm failed with status %d" % st)
        return (d[0], d[1]), (d[2], d[3], d[4])

    def close(self):
        if getattr(self, "handle", None):
            self.lib.refh_gas_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def ref_available():
    return os.path.exists(LIBREF)

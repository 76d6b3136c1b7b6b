"""ORACLE (test infrastructure only): CPU restatement of pyscf-forge's CSF
basis (pyscf/csf_fci/csdstring.py, csfstring.py, pyscf/lib/csf/csfstring.c).

* CSD order of determinants (csdstring.py:10-14, make_csd_mask :29-46,
  get_csdaddrs_shape :236-275): blocks by number of doubly occupied
  orbitals npair (ascending from max(0, na+nb-norb)), then the pair
  configuration (npair "electrons" in norb orbitals), the unpaired
  configuration (nspin electrons in the norb-npair remaining orbitals), and
  the spin configuration (nup up-spins among the nspin singles), each in
  pyscf's colex string order.
* Genealogical coupling strings (csf_gentable csfstring.py:806-849,
  FCICSFaddrs2str csfstring.c:241-293): bit = 1 couples up, 0 down, bits
  read from bit 0; address from the path-count table.
* Spin eigenvectors (FCICSFmakecsf csfstring.c:86-141): products of
  Clebsch-Gordan coefficients, with the fermion sign of commuting each
  down spin past the up spins before it.
* Block phase (-1)^ncomm, ncomm = npair(npair-1)/2 + npair*max(0, nb-npair)
  (csfstring.py:623-627).
* CSF diagonal (make_hdiag_csf csf.py:70-144, FCICSFhdiag csfstring.c:295):
  diag(U^T H_conf U) with H_conf the spin-coupling block of H inside one
  configuration.

Checked in tests/test_csf_oracle.py against the reference's own csfstring.c
(compiled into oracle/_ref) and against spin invariants.
"""

import math

import numpy

import oracle


def _colex_strings(n, k):
    return oracle.make_strings(n, k).astype(numpy.int64) if n >= 0 and 0 <= k <= n else numpy.zeros(0, numpy.int64)


def csd_shape(norb, na, nb):
    """(min_npair, offsets, dconf_size, sconf_size, spin_size) per npair block."""
    nless = min(na, nb)
    min_npair = max(0, na + nb - norb)
    rows = []
    off = 0
    for npair in range(min_npair, nless + 1):
        nspin = na + nb - 2 * npair
        nup = (nspin + na - nb) // 2
        d = math.comb(norb, npair)
        sc = math.comb(norb - npair, nspin)
        sp = math.comb(nspin, nup)
        rows.append((npair, off, d, sc, sp))
        off += d * sc * sp
    return min_npair, rows


def gentable(nspin, smult):
    """Genealogical path-count table (csfstring.py:806-849)."""
    if smult > nspin + 1:
        return numpy.zeros((1, 1), dtype=numpy.int64)
    n0 = (nspin - (smult - 1)) // 2
    n1 = (nspin + (smult - 1)) // 2
    g = numpy.zeros((n0 + 1, n1 + 1), dtype=numpy.int64)
    g[n0, n0:] = 1
    for i0 in range(n0, 0, -1):
        row = list(g[i0, i0:])
        row = [sum(row[i1:]) for i1 in range(len(row))]
        row = [row[0]] + row
        g[i0 - 1, i0 - 1:] = row
    return g


def count_csfs(nspin, smult):
    if (smult - 1) % 2 != nspin % 2 or smult > nspin + 1:
        return 0
    return int(gentable(nspin, smult)[0, 0])


def coupling_strings(nspin, smult):
    """All coupling strings in address order (addrs2str, csfstring.c:241-293)."""
    ncsf = count_csfs(nspin, smult)
    if ncsf == 0:
        return numpy.zeros(0, dtype=numpy.int64)
    g = gentable(nspin, smult)
    n1 = (nspin + smult - 1) // 2
    out = numpy.empty(ncsf, dtype=numpy.int64)
    for a in range(ncsf):
        s = 1
        rem = a
        i0 = i1 = 0
        for ispin in range(nspin):
            if i1 == n1:
                break
            if g[i0, i1 + 1] <= rem:
                rem -= g[i0, i1 + 1]
                i0 += 1
            else:
                s |= 1 << ispin
                i1 += 1
        out[a] = s
    return out


def spin_evecs(nspin, na, nb, smult):
    """umat (ndet_spin x ncsf): Clebsch-Gordan coupling (csfstring.c:86-141)."""
    nup = (nspin + na - nb) // 2
    dets = _colex_strings(nspin, nup)
    coups = coupling_strings(nspin, smult)
    umat = numpy.zeros((len(dets), len(coups)))
    for i, det in enumerate(dets):
        for k, cp in enumerate(coups):
            two_s = 1
            two_ms = 1 if det & 1 else -1
            num, den, sgn = 1.0, 1.0, 1
            osgn = -two_ms
            ok = True
            for ispin in range(1, nspin):
                sup = (cp >> ispin) & 1
                msup = (det >> ispin) & 1
                if msup:
                    two_ms += 1
                    osgn = -osgn
                else:
                    two_ms -= 1
                    sgn *= osgn
                num *= (two_s + two_ms + 1) if sup == msup else (two_s + 1 - two_ms)
                if num == 0:
                    ok = False
                    break
                den *= two_s + 1
                if msup and not sup:
                    sgn = -sgn
                num /= 2
                two_s += 1 if sup else -1
            umat[i, k] = sgn * math.sqrt(num / den) if ok else 0.0
    return umat


def block_sign(npair, nb):
    ncomm = npair * (npair - 1) // 2 + npair * max(0, nb - npair)
    return -1.0 if ncomm % 2 else 1.0


def _expand(bits, positions):
    """Place the bits of `bits` (bit i -> positions[i])."""
    out = 0
    for i, p in enumerate(positions):
        if (bits >> i) & 1:
            out |= 1 << p
    return out


def csd_mask(norb, na, nb):
    """det address (ia * nstrb + ib) of every CSD position, in CSD order."""
    nstrb = math.comb(norb, nb)
    _, rows = csd_shape(norb, na, nb)
    total = rows[-1][1] + rows[-1][2] * rows[-1][3] * rows[-1][4] if rows else 0
    out = numpy.empty(total, dtype=numpy.int64)
    for npair, off, dsz, ssz, psz in rows:
        nspin = na + nb - 2 * npair
        nup = (nspin + na - nb) // 2
        dstrs = _colex_strings(norb, npair)
        sstrs = _colex_strings(norb - npair, nspin)
        pstrs = _colex_strings(nspin, nup)
        idx = off
        for dstr in dstrs:
            free = [o for o in range(norb) if not (dstr >> o) & 1]
            for sstr in sstrs:
                singles = [free[i] for i in range(len(free)) if (sstr >> i) & 1]
                sbits = _expand(sstr, free)
                for pstr in pstrs:
                    up = _expand(pstr, singles)
                    a = int(dstr) | up
                    b = int(dstr) | (sbits & ~up)
                    out[idx] = oracle.str2addr(norb, na, a) * nstrb + oracle.str2addr(norb, nb, b)
                    idx += 1
    return out


def csf_shape(norb, na, nb, smult):
    """Per npair block: (npair, csd_offset, csf_offset, nconf, ndet_spin, ncsf)."""
    _, rows = csd_shape(norb, na, nb)
    out = []
    coff = 0
    for npair, off, dsz, ssz, psz in rows:
        nspin = na + nb - 2 * npair
        ncsf = count_csfs(nspin, smult) if abs(na - nb) <= smult - 1 else 0
        out.append((npair, off, coff, dsz * ssz, psz, ncsf))
        coff += dsz * ssz * ncsf
    return out, coff


def det2csf(c, norb, na, nb, smult, mask=None):
    c = numpy.asarray(c, dtype=numpy.float64).reshape(-1)
    mask = csd_mask(norb, na, nb) if mask is None else mask
    blocks, ncsf_all = csf_shape(norb, na, nb, smult)
    out = numpy.zeros(ncsf_all)
    for npair, off, coff, nconf, ndet, ncsf in blocks:
        if ncsf == 0:
            continue
        u = block_sign(npair, nb) * spin_evecs(na + nb - 2 * npair, na, nb, smult)
        g = c[mask[off:off + nconf * ndet]].reshape(nconf, ndet)
        out[coff:coff + nconf * ncsf] = (g @ u).reshape(-1)
    return out


def csf2det(x, norb, na, nb, smult, mask=None):
    mask = csd_mask(norb, na, nb) if mask is None else mask
    blocks, ncsf_all = csf_shape(norb, na, nb, smult)
    out = numpy.zeros(math.comb(norb, na) * math.comb(norb, nb))
    for npair, off, coff, nconf, ndet, ncsf in blocks:
        if ncsf == 0:
            continue
        u = block_sign(npair, nb) * spin_evecs(na + nb - 2 * npair, na, nb, smult)
        xs = numpy.asarray(x)[coff:coff + nconf * ncsf].reshape(nconf, ncsf)
        out[mask[off:off + nconf * ndet]] = (xs @ u.T).reshape(-1)
    return out


def hdiag_csf(h1e, eri, norb, na, nb, smult, mask=None):
    """Diagonal of H in the CSF basis (make_hdiag_csf, csf.py:70-144):
    per configuration, H_conf[x,y] = hdiag_det on the diagonal and
    sgn * (e0 e1|e1 e0) between spin determinants that differ by exchanging
    the spins of two singly occupied orbitals e0 < e1
    (FCICSFhdiag csfstring.c:295-361); hdiag_csf = diag(U^T H_conf U)."""
    mask = csd_mask(norb, na, nb) if mask is None else mask
    hdet = oracle.make_hdiag(h1e, eri, norb, (na, nb))
    eri4 = oracle.unpack_pair_matrix(oracle.pair_matrix(eri, norb), norb)
    nstrb = math.comb(norb, nb)
    blocks, ncsf_all = csf_shape(norb, na, nb, smult)
    out = numpy.zeros(ncsf_all)
    for npair, off, coff, nconf, ndet, ncsf in blocks:
        if ncsf == 0:
            continue
        u = block_sign(npair, nb) * spin_evecs(na + nb - 2 * npair, na, nb, smult)
        for conf in range(nconf):
            addrs = mask[off + conf * ndet: off + (conf + 1) * ndet]
            astr = [oracle.addr2str(norb, na, a // nstrb) for a in addrs]
            bstr = [oracle.addr2str(norb, nb, a % nstrb) for a in addrs]
            H = numpy.diag(hdet[addrs])
            for x in range(ndet):
                for y in range(x):
                    exc = astr[x] ^ astr[y]
                    somo = astr[x] ^ bstr[x]
                    if bin(exc).count("1") != 2:
                        continue
                    sgn, esgn, e = -1, 1, []
                    for o in range(norb):
                        if (somo >> o) & 1:
                            esgn = -esgn
                        if (exc >> o) & 1:
                            e.append(o)
                            sgn *= esgn
                    v = sgn * eri4[e[0], e[1], e[1], e[0]]
                    H[x, y] = H[y, x] = v
            out[coff + conf * ncsf: coff + (conf + 1) * ncsf] = numpy.einsum("dk,de,ek->k", u, H, u)
    return out

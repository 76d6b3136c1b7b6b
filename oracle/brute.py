"""ORACLE (test infrastructure only): Slater-Condon brute force.

Builds the full determinant-basis Hamiltonian of a small CAS problem from
second quantisation, independently of any link table or absorbed tensor:

    H = sum_pq h_pq E_pq + 1/2 sum_pqrs (pq|rs) (E_pq E_rs - delta_qr E_ps)

with E_pq = E^a_pq (x) 1 + 1 (x) E^b_pq and E^s_pq = a^+_ps a_qs built by
acting on bit strings directly.  This is the same idea as the brute-force H
of pyscf/csf_fci/test/test_csf.py:60-74 (contract_2e on unit vectors), done
here without contract_2e so that it checks contract_2e.  Determinant order is
(alpha address, beta address), addresses in colex order (pyscf cistring).
"""

import itertools

import numpy


def _strings(n, k):
    """All k-subsets of n orbitals as ints, ordered by value (colex order)."""
    out = [sum(1 << p for p in occ) for occ in itertools.combinations(range(n), k)]
    return sorted(out)


def one_body_ops(n, k):
    """E_pq on the k-electron string space: array (n, n, nstr, nstr)."""
    strs = _strings(n, k)
    index = {s: i for i, s in enumerate(strs)}
    E = numpy.zeros((n, n, len(strs), len(strs)))
    for j, s in enumerate(strs):
        for q in range(n):
            if not (s >> q) & 1:
                continue
            # a_q: sign = (-1)^(# occupied below q)
            sign_q = (-1) ** bin(s & ((1 << q) - 1)).count("1")
            s1 = s ^ (1 << q)
            for p in range(n):
                if (s1 >> p) & 1:
                    continue
                sign_p = (-1) ** bin(s1 & ((1 << p) - 1)).count("1")
                s2 = s1 | (1 << p)
                E[p, q, index[s2], j] += sign_p * sign_q
    return E


def hamiltonian(h1e, eri, n, nelec):
    """Dense H in the (nstr_a*nstr_b)^2 determinant basis.

    eri: (n,n,n,n) chemist tensor, or a 4-fold (npair,npair) pair matrix.
    """
    na, nb = nelec
    eri = numpy.asarray(eri, dtype=numpy.float64)
    if eri.ndim == 2:
        p, q = numpy.indices((n, n))
        pq = numpy.maximum(p, q) * (numpy.maximum(p, q) + 1) // 2 + numpy.minimum(p, q)
        eri = eri[pq[:, :, None, None], pq[None, None, :, :]]
    Ea = one_body_ops(n, na)
    Eb = one_body_ops(n, nb)
    Ia = numpy.eye(Ea.shape[2])
    Ib = numpy.eye(Eb.shape[2])
    N = Ea.shape[2] * Eb.shape[2]
    E = numpy.empty((n, n, N, N))
    for p in range(n):
        for q in range(n):
            E[p, q] = numpy.kron(Ea[p, q], Ib) + numpy.kron(Ia, Eb[p, q])
    H = numpy.einsum("pq,pqij->ij", h1e, E)
    # 1/2 sum (pq|rs) E_pq E_rs - 1/2 sum_{pqs} (pq|qs) E_ps
    A = numpy.einsum("pqrs,pqij->rsij", eri, E)
    for r in range(n):
        for s in range(n):
            H += 0.5 * A[r, s] @ E[r, s]
    K = numpy.einsum("pqqs->ps", eri)
    H -= 0.5 * numpy.einsum("ps,psij->ij", K, E)
    return H


def spin_square_matrix(n, nelec):
    """S^2 in the (alpha address, beta address) determinant basis:
    S^2 = S_- S_+ + S_z (S_z + 1), S_+ = sum_p a^+_{p alpha} a_{p beta},
    with alpha spin-orbitals ordered before beta ones (the sign convention of
    the alpha-string x beta-string product basis)."""
    na, nb = nelec
    sa, sb = _strings(n, na), _strings(n, nb)
    ia = {s: i for i, s in enumerate(sa)}
    ib = {s: j for j, s in enumerate(sb)}
    N = len(sa) * len(sb)
    S2 = numpy.zeros((N, N))
    sz = 0.5 * (na - nb)
    for i, a in enumerate(sa):
        for j, b in enumerate(sb):
            col = i * len(sb) + j
            S2[col, col] += sz * (sz + 1)
            # S_- S_+ |a b> = sum_{p,q} a+_{q beta} a_{q alpha} a+_{p alpha} a_{p beta} |a b>
            for p in range(n):
                if not (b >> p) & 1 or (a >> p) & 1:
                    continue
                # a_{p beta}: sign = (-1)^(na + #beta below p) in the combined ordering
                s1 = (-1) ** (na + bin(b & ((1 << p) - 1)).count("1"))
                b1 = b ^ (1 << p)
                s2 = (-1) ** bin(a & ((1 << p) - 1)).count("1")
                a1 = a | (1 << p)
                for q in range(n):
                    if not (a1 >> q) & 1 or (b1 >> q) & 1:
                        continue
                    s3 = (-1) ** bin(a1 & ((1 << q) - 1)).count("1")
                    a2 = a1 ^ (1 << q)
                    s4 = (-1) ** (na + bin(b1 & ((1 << q) - 1)).count("1"))
                    b2 = b1 | (1 << q)
                    if a2 not in ia:
                        continue
                    row = ia[a2] * len(sb) + ib[b2]
                    S2[row, col] += s1 * s2 * s3 * s4
    return S2


def det_spin_square_matrix(stra, strb, nelec):
    """S^2 over an explicit determinant list (stra[i], strb[i]), restricted
    to the list (P S^2 P), by the ladder construction of spin_square_matrix
    (S_- S_+ + S_z (S_z + 1), alpha spin-orbitals before beta ones); strings
    of up to 64 orbitals.  Small lists (dense output)."""
    na, nb = nelec
    sa = [int(x) for x in numpy.asarray(stra, dtype=numpy.uint64)]
    sb = [int(x) for x in numpy.asarray(strb, dtype=numpy.uint64)]
    index = {(a, b): i for i, (a, b) in enumerate(zip(sa, sb))}
    N = len(sa)
    S2 = numpy.zeros((N, N))
    sz = 0.5 * (na - nb)
    pc = lambda x: bin(x).count("1")
    for col, (a, b) in enumerate(zip(sa, sb)):
        S2[col, col] += sz * (sz + 1)
        bb = b & ~a
        while bb:
            p = (bb & -bb).bit_length() - 1
            bb &= bb - 1
            s1 = (-1) ** (na + pc(b & ((1 << p) - 1)))
            b1 = b ^ (1 << p)
            s2 = (-1) ** pc(a & ((1 << p) - 1))
            a1 = a | (1 << p)
            aa = a1 & ~b1
            while aa:
                q = (aa & -aa).bit_length() - 1
                aa &= aa - 1
                s3 = (-1) ** pc(a1 & ((1 << q) - 1))
                a2 = a1 ^ (1 << q)
                s4 = (-1) ** (na + pc(b1 & ((1 << q) - 1)))
                row = index.get((a2, b1 | (1 << q)))
                if row is not None:
                    S2[row, col] += s1 * s2 * s3 * s4
    return S2

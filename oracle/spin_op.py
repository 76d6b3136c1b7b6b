"""ORACLE (test infrastructure only): spin ladder operators on determinant
vectors, restating pyscf/csf_fci/spin_op.py:22-71 (contract_sladder) with
plain loops over bit strings.

S- = sum_p a+_{p beta} a_{p alpha} and S+ = sum_p a+_{p alpha} a_{p beta}
act on |Ia Ib> (alpha operators before beta ones).  The single-operator
signs are those of pyscf's cistring (gen_des_str_index /
gen_cre_str_index, a third-party routine absent here): a creation or
annihilation at p counts the occupied orbitals above p; moving the beta
operator past the alpha string gives (-1)^(na - 1) for S- and (-1)^na for
S+ (spin_op.py:54-56).  The sign convention is pinned in
tests/test_spin_op_oracle.py by the CSF basis (pinned to the reference's
csfstring.c): S- maps each CSF |S, M> to +norm_sdown |S, M - 1> with the
same coupling (the Condon-Shortley phase); the magnitudes by
S+ S- + Sz(Sz - 1) = S^2 against the brute-force S^2.
"""

import numpy

import oracle


def contract_sladder(fcivec, norb, nelec, op=-1):
    na, nb = nelec
    if (op == -1 and (na == 0 or nb == norb)) or (op == 1 and (na == norb or nb == 0)):
        return numpy.zeros((0, 0))
    sa = oracle.make_strings(norb, na)
    sb = oracle.make_strings(norb, nb)
    da = {int(s): i for i, s in enumerate(oracle.make_strings(norb, na + op))}
    db = {int(s): i for i, s in enumerate(oracle.make_strings(norb, nb - op))}
    c = numpy.asarray(fcivec, dtype=numpy.float64).reshape(len(sa), len(sb))
    out = numpy.zeros((len(da), len(db)))
    fac = (-1) ** (na - 1 if op == -1 else na)
    for a, A in enumerate(int(x) for x in sa):
        for b, B in enumerate(int(x) for x in sb):
            if c[a, b] == 0.0:
                continue
            for p in range(norb):
                bit = 1 << p
                if op == -1 and ((A & bit) == 0 or (B & bit) != 0):
                    continue
                if op == 1 and ((A & bit) != 0 or (B & bit) == 0):
                    continue
                sign = (-1) ** (bin(A >> (p + 1)).count("1") + bin(B >> (p + 1)).count("1"))
                out[da[A ^ bit], db[B ^ bit]] += fac * sign * c[a, b]
    return out


def norm_sdown(smult, nelec):
    s2, m2 = smult - 1, nelec[0] - nelec[1]
    return numpy.sqrt((s2 - m2 + 2) * (s2 + m2) / 4)


def norm_sup(smult, nelec):
    s2, m2 = smult - 1, nelec[0] - nelec[1]
    return numpy.sqrt((s2 - m2) * (s2 + m2 + 2) / 4)

"""Synthetic CAS problems for the sigma benchmarks and parity cases.

The distributions are those of the benchmark definitions; the generator and
the order of the draws are fixed here (numpy.random.default_rng(seed)):

1. ``A ~ N(0, 0.1)`` of shape (n, n);
   ``h1e = (A + A^T)/2 + diag(linspace(-2, -0.5, n))``
2. ``B ~ N(0, 0.05)`` of shape (npair, npair); ``eri = B B^T``
   (an 8-fold symmetric, positive semi-definite pair matrix)
3. ``c ~ N(0, 1)`` of shape (nstr_alpha, nstr_beta), normalised.

``make_integrals`` is the closed-form recipe of the reference's GAS tests
(pyscf/mcscf/test/test_gasci.py:41-58), used for fixture cases.
"""

import math

import numpy


def num_strings(n, k):
    return math.comb(n, k)


def make_problem(norb, nelec, seed=0, with_ci=True):
    na, nb = nelec
    rng = numpy.random.default_rng(seed)
    a = rng.normal(0.0, 0.1, (norb, norb))
    h1e = (a + a.T) * 0.5 + numpy.diag(numpy.linspace(-2.0, -0.5, norb))
    npair = norb * (norb + 1) // 2
    b = rng.normal(0.0, 0.05, (npair, npair))
    eri = b @ b.T
    eri = (eri + eri.T) * 0.5
    if not with_ci:
        return h1e, eri
    c = rng.normal(0.0, 1.0, (num_strings(norb, na), num_strings(norb, nb)))
    c /= numpy.linalg.norm(c)
    return h1e, eri, c


def make_integrals(norb):
    """Closed-form integrals of pyscf/mcscf/test/test_gasci.py:41-58."""
    h1e = numpy.empty((norb, norb))
    for p in range(norb):
        for q in range(norb):
            h1e[p, q] = 0.07 * (p + q + 2) + 0.013 * ((p + 1) * (q + 3) % 7)
    h1e = (h1e + h1e.T) * 0.5
    npair = norb * (norb + 1) // 2
    eri = numpy.empty((npair, npair))
    for pq in range(npair):
        for rs in range(npair):
            lo, hi = min(pq, rs), max(pq, rs)
            eri[pq, rs] = 0.011 * (lo + 1) + 0.003 * (hi + 2)
    return h1e, eri


def batch_manifest(nproblems=64, nroots=4):
    """Per-problem (norb, (na, nb), seed) draws of the batched workload:
    n ~ U{6..12} from default_rng(i), nb = n//2, na = nb + d, d in {0, 1}."""
    out = []
    for i in range(nproblems):
        rng = numpy.random.default_rng(i)
        n = int(rng.integers(6, 13))
        d = int(rng.integers(0, 2))
        nb = n // 2
        out.append((n, (nb + d, nb), i))
    return out

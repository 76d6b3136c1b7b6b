"""The oracle is checked here against (i) the reference's libfci_gas golden
vectors (tests/golden, written by tests/golden/gen_golden.py), (ii) a
from-scratch Slater-Condon brute force and (iii) closed forms / invariants.
CPU only."""

import itertools
import math

import numpy
import pytest

import oracle
from oracle import brute
from conftest import golden_cases, load_golden, package


syn = package().synthetic


@pytest.mark.parametrize("n,k", [(1, 1), (4, 2), (8, 4), (9, 3), (12, 6), (16, 8), (16, 1)])
def test_addressing_roundtrip(n, k):
    strs = oracle.make_strings(n, k)
    assert strs.size == math.comb(n, k)
    # ascending integer (colex) order, every string has k bits
    assert numpy.all(numpy.diff(strs.astype(numpy.int64)) > 0)
    assert all(bin(int(s)).count("1") == k for s in strs[:: max(1, strs.size // 500)])
    expect = sorted(sum(1 << p for p in c) for c in itertools.combinations(range(n), k))
    if n <= 12:
        assert [int(s) for s in strs] == expect
    for a in range(0, strs.size, max(1, strs.size // 300)):
        assert oracle.str2addr(n, k, strs[a]) == a


@pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built")
def test_addressing_matches_reference():
    ref = oracle.RefLib()
    for n, (na, nb) in [(8, (4, 4)), (10, (3, 5))]:
        for spin, k in ((0, na), (1, nb)):
            for a, s in enumerate(oracle.make_strings(n, k)):
                assert ref.str2addr(n, (na, nb), spin, s) == a


def _independent_link(n, s, p, q):
    """a_p^+ a_q |s> by explicit operator action (sign from occupations below)."""
    if not (s >> q) & 1:
        return None
    sq = (-1) ** bin(s & ((1 << q) - 1)).count("1")
    s1 = s ^ (1 << q)
    if (s1 >> p) & 1:
        return None
    sp = (-1) ** bin(s1 & ((1 << p) - 1)).count("1")
    return s1 | (1 << p), sp * sq


@pytest.mark.parametrize("n,k", [(6, 3), (8, 4), (7, 2)])
def test_link_table_closed_form(n, k):
    link = oracle.gen_linkstr_index(n, k)
    strs = oracle.make_strings(n, k)
    for a, s in enumerate(strs):
        s = int(s)
        occ = [p for p in range(n) if (s >> p) & 1]
        vir = [p for p in range(n) if not (s >> p) & 1]
        order = [(i, i) for i in occ] + [(v, i) for i in occ for v in vir]
        assert [tuple(r[:2]) for r in link[a]] == order
        for cre, des, addr, sign in link[a]:
            s1, sg = _independent_link(n, s, cre, des)
            assert strs[addr] == s1 and sign == sg
    tl = oracle.gen_linkstr_index(n, k, tril=True)
    assert numpy.array_equal(tl[:, :, 2:], link[:, :, 2:])
    p, q = link[:, :, 0], link[:, :, 1]
    assert numpy.array_equal(tl[:, :, 0], numpy.maximum(p, q) * (numpy.maximum(p, q) + 1) // 2 + numpy.minimum(p, q))


@pytest.mark.parametrize("name", golden_cases())
def test_link_table_rowsets_vs_reference(name):
    g = load_golden(name)
    n, (na, nb) = g["norb"], g["nelec"]
    for k, tab in ((na, g["linka"]), (nb, g["linkb"])):
        mine = oracle.gen_linkstr_index(n, k)
        assert mine.shape == tab.shape
        for r in range(mine.shape[0]):
            assert sorted(map(tuple, mine[r])) == sorted(map(tuple, tab[r]))


@pytest.mark.parametrize("name", golden_cases())
def test_sigma_hdiag_vs_reference(name):
    g = load_golden(name)
    n, ne = g["norb"], g["nelec"]
    h2e = oracle.absorb_h1e(g["h1e"], g["eri"], n, ne, 0.5)
    numpy.testing.assert_allclose(h2e, g["h2e"], rtol=0, atol=1e-14)
    sigma = oracle.contract_2e(h2e, g["ci"], n, ne)
    scale = numpy.abs(g["sigma"]).max()
    numpy.testing.assert_allclose(sigma, g["sigma"], rtol=0, atol=1e-13 * scale)
    hd = oracle.make_hdiag(g["h1e"], g["eri"], n, ne).reshape(g["hdiag"].shape)
    numpy.testing.assert_allclose(hd, g["hdiag"], rtol=0, atol=1e-12)


@pytest.mark.parametrize("name", golden_cases())
def test_rdm_vs_reference(name):
    g = load_golden(name)
    n, ne = g["norb"], g["nelec"]
    (d1a, d1b), (aa, ab, bb) = oracle.trans_rdm12s(g["ci"], g["ket"], n, ne)
    for mine, key in ((d1a, "dm1a"), (d1b, "dm1b"), (aa, "dm2aa"), (ab, "dm2ab"), (bb, "dm2bb")):
        numpy.testing.assert_allclose(mine, g[key], rtol=0, atol=1e-12)


@pytest.mark.parametrize("n,ne,seed", [(4, (2, 2), 1), (5, (3, 2), 2), (6, (3, 3), 3), (6, (2, 1), 4), (5, (5, 0), 5),
                                       # degenerate spaces: one orbital, empty and full shells,
                                       # one spin without electrons
                                       (1, (1, 0), 6), (1, (1, 1), 7), (2, (1, 1), 8), (4, (0, 0), 9),
                                       (4, (4, 4), 10), (6, (0, 3), 11), (5, (5, 2), 12), (7, (1, 1), 13)])
def test_sigma_absorb_vs_brute_force(n, ne, seed):
    h1e, eri, c = syn.make_problem(n, ne, seed=seed)
    H = brute.hamiltonian(h1e, eri, n, ne)
    h2e = oracle.absorb_h1e(h1e, eri, n, ne, 0.5)
    sig = oracle.contract_2e(h2e, c, n, ne).reshape(-1)
    numpy.testing.assert_allclose(sig, H @ c.reshape(-1), rtol=0, atol=1e-13)
    numpy.testing.assert_allclose(oracle.make_hdiag(h1e, eri, n, ne), numpy.diag(H), atol=1e-13)
    # the hermiticity of the brute force H itself
    numpy.testing.assert_allclose(H, H.T, atol=1e-13)


def test_rdm_identities_and_energy():
    n, ne = 7, (4, 3)
    h1e, eri, c = syn.make_problem(n, ne, seed=9)
    N = sum(ne)
    dm1, dm2 = oracle.make_rdm12(c, n, ne)
    assert abs(numpy.trace(dm1) - N) < 1e-12
    numpy.testing.assert_allclose(numpy.einsum("pqrr->pq", dm2), (N - 1) * dm1, atol=1e-12)
    e_rdm = numpy.einsum("pq,pq", h1e, dm1) + 0.5 * numpy.einsum(
        "pqrs,pqrs", oracle.unpack_pair_matrix(eri, n), dm2)
    assert abs(e_rdm - oracle.energy(h1e, eri, c, n, ne)) < 1e-12


def test_planted_sign_error_is_caught():
    """A wrong fermionic sign in the sigma (planted on the CPU, in a copy of
    the oracle's output path) must fail the brute-force comparison."""
    n, ne = 5, (3, 2)
    h1e, eri, c = syn.make_problem(n, ne, seed=11)
    H = brute.hamiltonian(h1e, eri, n, ne)
    # flip the sign of every odd link: emulate with a sign-flipped alpha string
    flip = numpy.ones(c.shape[0])
    flip[1::2] = -1
    h2e = oracle.absorb_h1e(h1e, eri, n, ne, 0.5)
    bad = (flip[:, None] * oracle.contract_2e(h2e, flip[:, None] * c * flip[:, None] ** 2, n, ne)).reshape(-1)
    bad_sigma = oracle.contract_2e(h2e, c * flip[:, None], n, ne).reshape(-1)
    assert numpy.abs(bad_sigma - H @ c.reshape(-1)).max() > 1e-3
    assert numpy.isfinite(bad).all()


def test_lowest_eigenvalue_matches_dense():
    n, ne = 6, (3, 3)
    h1e, eri = syn.make_problem(n, ne, seed=5, with_ci=False)
    w = numpy.linalg.eigvalsh(brute.hamiltonian(h1e, eri, n, ne))
    e, v = oracle.lowest_eigenvalues(h1e, eri, n, ne, nroots=2)
    numpy.testing.assert_allclose(e, w[:2], atol=1e-10)


# --- Slater-Condon rows (orc_sigma_rows): the spot-check oracle for spaces
# too large for a full CPU sigma.  Pinned against the brute-force H (which
# builds H from explicit operator products, no Slater-Condon rules) and
# against the reference's libfci_gas sigma vectors.

@pytest.mark.parametrize("n,ne,seed", [(4, (2, 2), 0), (5, (3, 1), 1), (6, (3, 3), 2),
                                       (6, (0, 2), 3), (5, (2, 0), 4), (6, (4, 2), 5)])
def test_sigma_rows_vs_brute_force(n, ne, seed):
    h1e, eri, c = syn.make_problem(n, ne, seed=seed)
    H = brute.hamiltonian(h1e, eri, n, ne)
    dets = numpy.arange(H.shape[0])
    got = oracle.sigma_rows(h1e, eri, c, n, ne, dets)
    numpy.testing.assert_allclose(got, H @ c.reshape(-1), rtol=0, atol=1e-13)
    # unit vectors give whole rows of H
    for j in (0, H.shape[0] // 2, H.shape[0] - 1):
        e = numpy.zeros(H.shape[0])
        e[j] = 1.0
        numpy.testing.assert_allclose(oracle.sigma_rows(h1e, eri, e, n, ne, dets), H[:, j],
                                      rtol=0, atol=1e-14)


@pytest.mark.parametrize("name", golden_cases())
def test_sigma_rows_vs_golden(name):
    g = load_golden(name)
    n, ne = g["norb"], g["nelec"]
    ci = g["ci"].reshape(-1)
    rng = numpy.random.default_rng(11)
    dets = numpy.unique(numpy.concatenate([[0, ci.size - 1],
                                           rng.integers(0, ci.size, min(ci.size, 200))]))
    got = oracle.sigma_rows(g["h1e"], g["eri"], ci, n, ne, dets)
    ref = g["sigma"].reshape(-1)[dets]
    numpy.testing.assert_allclose(got, ref, rtol=0, atol=1e-13 * numpy.abs(g["sigma"]).max())


def test_sigma_rows_vs_contract_2e_12_orbitals():
    n, ne = 12, (6, 5)
    h1e, eri, c = syn.make_problem(n, ne, seed=9)
    ref = oracle.contract_2e(oracle.absorb_h1e(h1e, eri, n, ne, 0.5), c, n, ne).reshape(-1)
    dets = numpy.random.default_rng(3).integers(0, ref.size, 300)
    got = oracle.sigma_rows(h1e, eri, c, n, ne, dets)
    numpy.testing.assert_allclose(got, ref[dets], rtol=0, atol=1e-13 * numpy.abs(ref).max())


def test_sigma_rows_rejects_bad_index():
    h1e, eri, c = syn.make_problem(4, (2, 2), seed=0)
    with pytest.raises(ValueError):
        oracle.sigma_rows(h1e, eri, c, 4, (2, 2), [c.size])


# --- initial guess and preconditioner of the reference's solver

def test_splitmix64_known_answers():
    """The first four outputs of SplitMix64 from state 0 (Vigna's
    splitmix64.c): state k*golden mixed; the init-guess noise of root r at
    index 0 is mix((r+1)*golden), i.e. these values."""
    want = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F, 0xF88BB8A8724C81EC]
    got = [oracle.splitmix64_mix(k * 0x9E3779B97F4A7C15) for k in (1, 2, 3, 4)]
    assert got == want
    _, g = oracle.init_guess(numpy.arange(5.0), 4, noise_norm=1.0)
    for r, w in enumerate(want):
        assert numpy.sign(g[r][0] - (1.0 if r == 0 else 0.0)) == (1.0 if w & 1 else -1.0)


def test_init_guess_oracle_structure():
    rng = numpy.random.default_rng(4)
    hd = rng.normal(size=1000)
    addrs, g = oracle.init_guess(hd, 3)
    assert list(addrs) == list(numpy.argsort(hd)[:3])
    for r, (a, v) in enumerate(zip(addrs, g)):
        noise = v.copy()
        noise[a] -= 1.0
        assert abs(numpy.linalg.norm(noise) - 1e-3) < 1e-15
        # (1 + x) - 1 recovers x only to the rounding of 1
        assert numpy.allclose(numpy.abs(noise), 1e-3 / numpy.sqrt(1000.0), rtol=0, atol=2e-16)
    # the noise patterns of different roots are independent (not shifts)
    s0, s1 = numpy.sign(g[0]), numpy.sign(g[1])
    assert 0.3 < numpy.mean(s0[10:] == s1[10:]) < 0.7
    # whole range when the space is not larger than nroots
    a2, _ = oracle.init_guess([3.0, 1.0], 2)
    assert list(a2) == [0, 1]


def test_precond_oracle_level_shift():
    hd = numpy.array([1.0, 0.5 + 1e-4, 0.5 - 1e-4, -2.0, 0.5 + 1e-3])
    dx = numpy.ones(5)
    out = oracle.precond(dx, hd, 0.5, 1e-3)
    # small denominators of either sign become +level_shift
    numpy.testing.assert_allclose(out, [2.0, 1e3, 1e3, -0.4, 1.0 / (1e-3)], rtol=1e-12)

"""CPU checks of the restricted-space (GAS) oracle and of the product's
restriction front end, without a GPU.

The oracle's layout (sector ids and string counts, block offsets, sector
strings) must equal the reference libfci_gas's exactly, and its sigma,
diagonal and RDMs must equal libfci_gas's to rounding, on the fixtures
tests/golden/gas/*.npz that libfci_gas wrote (gen_gas_golden.py); when the
reference build oracle/_ref is present, also on extra random restrictions.
The restricted eigenvalues of the smallest cases are checked against a
Slater-Condon H restricted to the GAS determinants (oracle/brute.py), which
does not use the projection identity's sigma at all.

The front end (addons_gas) is checked against the reference tests' known
answers (test_gasci.py:84-163) and against the brute-force block sets of the
fixtures, which gen_gas_golden.py enumerates from the restrictions'
definitions.
"""

import glob
import json
import os

import numpy
import pytest

import oracle
from oracle import brute
from oracle import gas as og
from conftest import GOLDEN, package

pkg = package()
ag = pkg.addons_gas

CASES = sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN, "gas", "*.npz")))


def load(name):
    with numpy.load(os.path.join(GOLDEN, "gas", name + ".npz")) as f:
        d = {k: f[k] for k in f.files}
    d["gas_orbs"] = tuple(int(x) for x in d["gas_orbs"])
    d["nelec"] = tuple(int(x) for x in d["nelec"])
    d["restr_type"] = str(d["restr_type"])
    d["restr"] = json.loads(str(d["restr"]))
    return d


def test_fixture_set_present():
    assert len(CASES) >= 7


@pytest.mark.parametrize("name", CASES)
def test_oracle_layout_matches_libfci_gas(name):
    d = load(name)
    L = og.GasLayout(d["gas_orbs"], d["nelec"], d["blocks"])
    assert L.ndet == d["ci"].size
    assert L.sectors == [tuple(int(v) for v in r) for r in d["sector_occ"]]
    assert L.sector_nstr == [int(x) for x in d["sector_nstr"]]
    got = [(b["offset"], b["sa"], b["sb"], b["na"], b["nb"]) for b in L.block_desc]
    ref = list(zip(*(d["block_" + k].tolist() for k in ("offset", "sa", "sb", "na", "nb"))))
    assert got == ref
    strings = numpy.concatenate([L.sector_strings(s) for s in range(len(L.sectors))])
    numpy.testing.assert_array_equal(strings, d["sector_strings"])


@pytest.mark.parametrize("name", CASES)
def test_oracle_sigma_hdiag_rdm_match_libfci_gas(name):
    d = load(name)
    L = og.GasLayout(d["gas_orbs"], d["nelec"], d["blocks"])
    s = og.gas_contract_2e(d["h2e"], d["ci"], L)
    assert numpy.abs(s - d["sigma"]).max() < 1e-12 * max(1.0, numpy.abs(d["sigma"]).max())
    numpy.testing.assert_allclose(og.gas_make_hdiag(d["h1e"], d["eri"], L), d["hdiag"],
                                  rtol=0, atol=1e-11)
    (a, b), (aa, ab, bb) = og.gas_trans_rdm12s(d["ci"], d["ket"], L)
    for got, key in zip((a, b, aa, ab, bb), ("dm1a", "dm1b", "dm2aa", "dm2ab", "dm2bb")):
        numpy.testing.assert_allclose(got, d[key], rtol=0, atol=1e-12)


@pytest.mark.parametrize("name", [c for c in CASES if c.endswith(("n3", "n4"))])
def test_restricted_spectrum_vs_slater_condon(name):
    """Eigenvalues of the Slater-Condon H restricted to the GAS determinants
    equal those of libfci_gas's dense H and of the oracle's projection."""
    d = load(name)
    L = og.GasLayout(d["gas_orbs"], d["nelec"], d["blocks"])
    n = L.norb_total
    idx = L.cas_index()
    Hfull = brute.hamiltonian(d["h1e"], d["eri"], n, d["nelec"])
    w_sc = numpy.linalg.eigvalsh(Hfull[numpy.ix_(idx, idx)])[:4]
    numpy.testing.assert_allclose(w_sc, d["evals"][:len(w_sc)], rtol=0, atol=1e-11)
    w_or = numpy.linalg.eigvalsh(og.gas_hamiltonian(d["h2e"], L))[:4]
    numpy.testing.assert_allclose(w_or, w_sc, rtol=0, atol=1e-11)


def test_projection_identities():
    d = load("gas_sg_n8")
    L = og.GasLayout(d["gas_orbs"], d["nelec"], d["blocks"])
    c = d["ci"]
    full = og.gas2fci(c, L)
    numpy.testing.assert_array_equal(og.fci2gas(full, L), c)
    assert numpy.count_nonzero(full) == numpy.count_nonzero(c)
    # idempotent projection
    rng = numpy.random.default_rng(3)
    x = rng.normal(size=L.cas_shape)
    p1 = og.gas2fci(og.fci2gas(x, L), L)
    numpy.testing.assert_array_equal(og.gas2fci(og.fci2gas(p1, L), L), p1)


@pytest.mark.skipif(not oracle.ref_available(), reason="reference build absent")
def test_oracle_vs_libfci_gas_random_restrictions():
    rng = numpy.random.default_rng(11)
    ref = oracle.RefLib()
    for orbs, nelec, G in [((2, 3, 2), (3, 2), [[3, 2, 0], [2, 2, 1], [1, 3, 1], [2, 3, 0]]),
                           ((4, 2), (3, 3), [[5, 1], [4, 2], [6, 0]]),
                           ((1, 1, 2, 2), (2, 2), [[2, 1, 1, 0], [1, 1, 1, 1], [2, 2, 0, 0]])]:
        D = ag.blocks_from_supergroups(orbs, nelec, G)
        L = og.GasLayout(orbs, nelec, D)
        R = ref.gas(orbs, nelec, D)
        assert L.sectors == R.sectors and L.block_desc == R.block_desc
        n = sum(orbs)
        h1e, eri = pkg.synthetic.make_problem(n, nelec, seed=int(rng.integers(100)),
                                              with_ci=False)
        h2e = oracle.absorb_h1e(h1e, eri, n, nelec, 0.5)
        c = rng.normal(size=L.ndet)
        numpy.testing.assert_allclose(og.gas_contract_2e(h2e, c, L), R.contract_2e(h2e, c),
                                      rtol=0, atol=1e-12)


# ---------------------------------------------------------------- front end

@pytest.mark.parametrize("name", CASES)
def test_front_end_blocks_match_brute_force(name):
    d = load(name)
    orbs, D = ag.normalize_gas_spec(d["gas_orbs"], d["nelec"], d["restr"], d["restr_type"])
    assert tuple(orbs) == d["gas_orbs"]
    numpy.testing.assert_array_equal(D, d["blocks"])
    info = ag.check_kernel_limits(orbs, d["nelec"], D)
    assert info["ndet_estimate"] == d["ci"].size


def test_front_end_known_answers():
    expected = numpy.asarray([[0, 2, 2, 0], [1, 1, 1, 1], [2, 0, 0, 2]])
    explicit = [[2, 0, 0, 2], [1, 1, 1, 1], [0, 2, 2, 0], [1, 1, 1, 1]]
    for kind, r in (("spin-supergroup", explicit), ("supergroup", [[2, 2]]),
                    ("cumulative-occ", [[2, 2], [4, 4]])):
        numpy.testing.assert_array_equal(ag.normalize_gas_restr((2, 2), (2, 2), r, kind),
                                         expected)
    assert ag.is_spin_complete((2, 2), (2, 2), expected)
    assert not ag.is_spin_complete((2, 2), (2, 2), expected[:2])
    numpy.testing.assert_array_equal(
        ag.normalize_gas_restr(numpy.asarray([2, 2], dtype=numpy.int64), (2, 2),
                               numpy.asarray([[2, 2]], dtype=numpy.int64), "supergroup"),
        expected)
    ko, ras = ag.normalize_gas_spec((1, 2, 1), (2, 2), {"max_holes": 1, "max_particles": 1},
                                    "ras")
    cum = ag.normalize_gas_restr((1, 2, 1), (2, 2), [[1, 2], [3, 4], [4, 4]], "cumulative-occ")
    assert ko == (1, 2, 1)
    numpy.testing.assert_array_equal(ras, cum)
    # RAS with an empty RAS3 drops the empty space from the kernel space
    ko, D = ag.normalize_gas_spec((2, 2, 0), (2, 2), {"max_holes": 2, "max_particles": 0},
                                  "ras")
    assert ko == (2, 2)
    assert ag.check_kernel_limits((4,), (3, 1), [[3, 1]])["ndet_estimate"] == 16


@pytest.mark.parametrize("gas_orbs,restr,kind", [
    ((2.0, 2), [[2, 2]], "supergroup"),
    ((2, 2), [[2.0, 2]], "supergroup"),
    ((2, 2), [[2, 2], [4, 4.0]], "cumulative-occ"),
    ((1, 2, 1), {"max_holes": 1.0, "max_particles": 1}, "ras"),
    ((2, True), [[2, 2]], "supergroup"),
])
def test_front_end_rejects_non_integers(gas_orbs, restr, kind):
    with pytest.raises(TypeError):
        ag.normalize_gas_restr(gas_orbs, (2, 2), restr, kind)


def test_front_end_errors():
    with pytest.raises(NotImplementedError):
        ag.normalize_gas_spec((2, 2), (2, 2), [[2, 2]], "occupation-list")
    with pytest.raises(ValueError):
        ag.normalize_gas_spec((2, 2), (2, 2), None, "supergroup")
    with pytest.raises(ValueError):
        ag.normalize_gas_spec((2, 2), (2, 2))           # multi-space needs D
    with pytest.raises(ValueError):
        ag.normalize_gas_restr((2, 2), (2, 2), [[3, 2]], "supergroup")   # wrong total
    with pytest.raises(ValueError):
        ag.normalize_gas_restr((2, 2), (2, 2), [[3, 2], [4, 4]], "cumulative-occ")
    with pytest.raises(ValueError):
        ag.normalize_gas_restr((2, 2), (2, 1), [[2, 0, 1, 1]], "spin-supergroup")
    with pytest.raises(ValueError):
        ag.normalize_gas_spec((1, 2, 1), (2, 2), {"max_holes": 1}, "ras")


# --- Slater-Condon rows over an explicit determinant list
# (orc_det_sigma_rows): the spot-check oracle of sector-native GAS spaces,
# whose orbital counts leave no CAS array.  Pinned against the reference's
# libfci_gas restricted sigma (every row of every fixture) and against the
# CAS Slater-Condon rows when the list is the whole CAS space.

@pytest.mark.parametrize("name", CASES)
def test_det_sigma_rows_vs_libfci_gas(name):
    d = load(name)
    L = og.GasLayout(d["gas_orbs"], d["nelec"], d["blocks"])
    sa, sb = L.det_strings()
    got = oracle.det_sigma_rows(d["h1e"], d["eri"], d["ci"], L.norb_total, sa, sb,
                                numpy.arange(L.ndet))
    assert numpy.abs(got - d["sigma"]).max() < 1e-12 * max(1.0, numpy.abs(d["sigma"]).max())


@pytest.mark.parametrize("n,ne,seed", [(6, (3, 3), 0), (7, (4, 2), 1), (9, (5, 4), 2)])
def test_det_sigma_rows_full_cas_equals_sigma_rows(n, ne, seed):
    from conftest import package as _pk
    syn = _pk().synthetic
    h1e, eri, c = syn.make_problem(n, ne, seed=seed)
    sa0 = numpy.array([oracle.addr2str(n, ne[0], a) for a in range(oracle.num_strings(n, ne[0]))],
                      dtype=numpy.uint64)
    sb0 = numpy.array([oracle.addr2str(n, ne[1], b) for b in range(oracle.num_strings(n, ne[1]))],
                      dtype=numpy.uint64)
    sa, sb = numpy.repeat(sa0, sb0.size), numpy.tile(sb0, sa0.size)
    # a shuffled list: the result does not depend on the list order
    perm = numpy.random.default_rng(seed).permutation(sa.size)
    rows = numpy.arange(0, sa.size, 7)
    got = oracle.det_sigma_rows(h1e, eri, c.reshape(-1)[perm], n, sa[perm], sb[perm],
                                numpy.argsort(perm)[rows])
    ref = oracle.sigma_rows(h1e, eri, c, n, ne, rows)
    numpy.testing.assert_allclose(got, ref, rtol=0, atol=1e-13 * numpy.abs(ref).max())


def test_det_sigma_rows_planted_sign_fails():
    """A wrong sign of one coefficient changes the rows that couple to it:
    the check distinguishes the restricted H from a perturbed one."""
    d = load("gas_ras_n10")
    L = og.GasLayout(d["gas_orbs"], d["nelec"], d["blocks"])
    sa, sb = L.det_strings()
    c = d["ci"].copy()
    k = int(numpy.argmax(numpy.abs(c)))
    c[k] = -c[k]
    got = oracle.det_sigma_rows(d["h1e"], d["eri"], c, L.norb_total, sa, sb, numpy.arange(L.ndet))
    assert numpy.abs(got - d["sigma"]).max() > 1e-6


@pytest.mark.parametrize("n,ne,seed", [(4, (2, 2), 0), (5, (3, 2), 1), (6, (2, 3), 2)])
def test_det_hamiltonian_full_cas_equals_brute_force(n, ne, seed):
    from conftest import package as _pk
    h1e, eri, _ = _pk().synthetic.make_problem(n, ne, seed=seed)
    sa0 = numpy.array([oracle.addr2str(n, ne[0], a) for a in range(oracle.num_strings(n, ne[0]))],
                      dtype=numpy.uint64)
    sb0 = numpy.array([oracle.addr2str(n, ne[1], b) for b in range(oracle.num_strings(n, ne[1]))],
                      dtype=numpy.uint64)
    H = oracle.det_hamiltonian(h1e, eri, n, numpy.repeat(sa0, sb0.size), numpy.tile(sb0, sa0.size))
    numpy.testing.assert_allclose(H, brute.hamiltonian(h1e, eri, n, ne), rtol=0, atol=1e-13)


@pytest.mark.parametrize("name", [c for c in CASES if c.endswith(("n3", "n4"))])
def test_det_hamiltonian_gas_spectrum_vs_libfci_gas(name):
    d = load(name)
    L = og.GasLayout(d["gas_orbs"], d["nelec"], d["blocks"])
    sa, sb = L.det_strings()
    w = numpy.linalg.eigvalsh(oracle.det_hamiltonian(d["h1e"], d["eri"], L.norb_total, sa, sb))
    numpy.testing.assert_allclose(w[:4], d["evals"][:min(4, w.size)], rtol=0, atol=1e-11)


@pytest.mark.parametrize("n,ne", [(4, (2, 2)), (5, (3, 1)), (6, (3, 3))])
def test_det_spin_square_full_cas_equals_spin_square_matrix(n, ne):
    sa0 = numpy.array([oracle.addr2str(n, ne[0], a) for a in range(oracle.num_strings(n, ne[0]))],
                      dtype=numpy.uint64)
    sb0 = numpy.array([oracle.addr2str(n, ne[1], b) for b in range(oracle.num_strings(n, ne[1]))],
                      dtype=numpy.uint64)
    S = brute.det_spin_square_matrix(numpy.repeat(sa0, sb0.size), numpy.tile(sb0, sa0.size), ne)
    numpy.testing.assert_allclose(S, brute.spin_square_matrix(n, ne), rtol=0, atol=1e-14)
    w = numpy.linalg.eigvalsh(S)
    s = numpy.sqrt(w + 0.25) - 0.5
    assert numpy.abs(s * 2 - numpy.round(s * 2)).max() < 1e-10


@pytest.mark.parametrize("name", ["gas_sg_n4", "gas_ras_n4", "gas_cum_n3", "gas_sg_n8"])
def test_det_spin_square_gas_is_projection(name):
    d = load(name)
    L = og.GasLayout(d["gas_orbs"], d["nelec"], d["blocks"])
    sa, sb = L.det_strings()
    S = brute.det_spin_square_matrix(sa, sb, d["nelec"])
    idx = L.cas_index()
    ref = brute.spin_square_matrix(L.norb_total, d["nelec"])[numpy.ix_(idx, idx)]
    numpy.testing.assert_allclose(S, ref, rtol=0, atol=1e-14)

"""GPU parity of the CSF point-group packing (fcis_csf_symmetry / _irrep /
_pack / _unpack) and of the symmetry-adapted CSF solver (the reference's
csf_symm.FCISolver, pyscf/csf_fci/csf_symm.py:25-113) against brute force:
the determinant irrep is the XOR of orbsym over the occupied alpha and beta
orbitals, so the spectrum of irrep w and spin S is that of H restricted to
the determinants of irrep w, then to the S(S+1) eigenspace of S^2
(oracle/brute.py).  Tolerances: packing is a gather (exact), transforms
1e-13, energies 1e-9."""

import itertools

import numpy
import pytest

import oracle.brute as brute
from conftest import has_gpu, package

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not has_gpu(), reason="needs a HIP device")]

pkg = package()
csf = pkg.csf_fci

ORBSYM6 = [0, 3, 1, 0, 2, 3]
ORBSYM7 = [0, 1, 0, 3, 2, 1, 0]


def _symm_problem(orbsym, seed):
    """Random integrals that respect orbsym: h_pq = 0 unless the irreps of
    p and q agree, (pq|rs) = 0 unless their XOR product is 0."""
    o = numpy.asarray(orbsym)
    n = o.size
    rng = numpy.random.default_rng(seed)
    h1 = rng.standard_normal((n, n)) * 0.3
    h1 = (h1 + h1.T) * (o[:, None] == o[None, :])
    g = rng.standard_normal((n, n, n, n)) * 0.1
    g = g + g.transpose(1, 0, 2, 3)
    g = g + g.transpose(0, 1, 3, 2)
    g = g + g.transpose(2, 3, 0, 1)
    g *= (o[:, None, None, None] ^ o[None, :, None, None] ^ o[None, None, :, None]
          ^ o[None, None, None, :]) == 0
    return h1, g


def _string_irreps(n, k, orbsym):
    out = []
    for s in brute._strings(n, k):
        w = 0
        for p in range(n):
            if (s >> p) & 1:
                w ^= orbsym[p]
        out.append(w)
    return numpy.asarray(out)


def _det_irreps(n, na, nb, orbsym):
    return (_string_irreps(n, na, orbsym)[:, None] ^ _string_irreps(n, nb, orbsym)[None, :]
            ).reshape(-1)


def _levels(h1, g, n, na, nb, sm, orbsym, w):
    """Brute force: H in the determinants of irrep w, then in the S(S+1)
    eigenspace of S^2 there."""
    H = brute.hamiltonian(h1, g, n, (na, nb))
    irr = _det_irreps(n, na, nb, orbsym)
    idx = numpy.nonzero(irr == w)[0]
    off = numpy.nonzero(irr != w)[0]
    # the integrals are symmetric, so H does not couple irreps
    assert abs(H[numpy.ix_(idx, off)]).max(initial=0.0) < 1e-13
    S2 = brute.spin_square_matrix(n, (na, nb))[numpy.ix_(idx, idx)]
    ws, v = numpy.linalg.eigh(0.5 * (S2 + S2.T))
    s = 0.5 * (sm - 1)
    P = v[:, abs(ws - s * (s + 1)) < 1e-8]
    return numpy.linalg.eigvalsh(P.T @ H[numpy.ix_(idx, idx)] @ P), idx


@pytest.mark.parametrize("n,na,nb,sm,orbsym", [(6, 3, 3, 1, ORBSYM6), (6, 3, 3, 3, ORBSYM6),
                                               (6, 4, 2, 3, ORBSYM6), (7, 4, 3, 2, ORBSYM7)])
def test_pack_is_the_irrep_gather(n, na, nb, sm, orbsym):
    import torch
    t = csf.CSFTransformer(n, na, nb, sm)
    irreps = numpy.array([t.csf_irrep(i, orbsym) for i in range(t.ncsf)])
    # every CSF's determinant support lies in the irrep of its configuration
    det_irr = _det_irreps(n, na, nb, orbsym)
    eye = torch.eye(t.ncsf, dtype=torch.float64, device="cuda")
    dets = torch.empty((t.ncsf, t.ndet), dtype=torch.float64, device="cuda")
    t.csf2det_dev(eye, dets, t.ncsf)
    dets = dets.cpu().numpy()
    for i in range(t.ncsf):
        nz = numpy.nonzero(abs(dets[i]) > 1e-14)[0]
        assert nz.size and (det_irr[nz] == irreps[i]).all(), i
    total = 0
    rng = numpy.random.default_rng(n + 7 * sm)
    x = rng.standard_normal((3, t.ncsf))
    for w in range(8):
        t._update_symm_cache(orbsym, w)
        mask = irreps == w
        assert t.nsym == int(mask.sum())
        total += t.nsym
        if t.nsym == 0:
            continue
        packed = t.pack_csf(x)
        numpy.testing.assert_array_equal(packed, x[:, mask])
        full = t.unpack_csf(packed)
        numpy.testing.assert_array_equal(full, numpy.where(mask, x, 0.0))
        # packed det <-> CSF transforms: the det vector of a packed CSF vector
        # has irrep w, and det2csf of it returns the packed vector
        y = rng.standard_normal((2, t.nsym))
        d = t.vec_csf2det(y, normalize=False)
        assert abs(d[:, det_irr != w]).max(initial=0.0) == 0.0
        numpy.testing.assert_allclose(t.vec_det2csf(d, normalize=False), y, rtol=0, atol=1e-13)
    assert total == t.ncsf_full
    assert t.ncsf == t.nsym
    t.close()


@pytest.mark.parametrize("n,na,nb,sm,orbsym", [(6, 3, 3, 1, ORBSYM6), (6, 3, 3, 3, ORBSYM6),
                                               (6, 4, 2, 3, ORBSYM6)])
def test_symm_solver_vs_brute_force(n, na, nb, sm, orbsym):
    h1, g = _symm_problem(orbsym, 5 * n + sm)
    t = csf.CSFTransformer(n, na, nb, sm)
    present = sorted({t.csf_irrep(i, orbsym) for i in range(t.ncsf)})
    t.close()
    for w in present:
        ref, idx = _levels(h1, g, n, na, nb, sm, orbsym, w)
        k = min(2, ref.size)
        sol = csf.FCISolver(smult=sm, orbsym=orbsym, wfnsym=w)
        e, ci = sol.kernel(h1, g, n, (na, nb), nroots=k)
        numpy.testing.assert_allclose(numpy.atleast_1d(e), ref[:k], rtol=0, atol=1e-9)
        assert sol.wfnsym == w and sol.orbsym == orbsym
        e2, ci2 = sol.kernel(h1, g, n, (na, nb), nroots=k, davidson_only=True, pspace_size=0)
        numpy.testing.assert_allclose(numpy.atleast_1d(e2), ref[:k], rtol=0, atol=1e-9)
        cis = ci2 if k > 1 else [ci2]
        for c in cis:
            flat = c.reshape(-1)
            out = numpy.setdiff1d(numpy.arange(flat.size), idx)
            assert abs(flat[out]).max(initial=0.0) < 1e-12
            _, mult = pkg.FCISolver().spin_square(c, n, (na, nb))
            assert abs(mult - sm) < 1e-6


def test_symm_solver_davidson_with_pspace_and_guess():
    """(7o; 4a,3b) doublet: Davidson in the packed space with the Olsen
    pspace preconditioner on a 12-CSF block; wfnsym defaults to the irrep
    of the aufbau determinant and kwargs apply to one call only."""
    n, na, nb, sm = 7, 4, 3, 2
    orbsym = ORBSYM7
    h1, g = _symm_problem(orbsym, 77)
    aufbau = orbsym[3]
    ref, _ = _levels(h1, g, n, na, nb, sm, orbsym, aufbau)
    sol = csf.FCISolver(smult=sm)
    e, ci = sol.kernel(h1, g, n, (na, nb), orbsym=orbsym, davidson_only=True, pspace_size=12,
                       nroots=3)
    assert sol.orbsym is None and sol.wfnsym is None
    numpy.testing.assert_allclose(e, ref[:3], rtol=0, atol=1e-9)
    assert sol.guess_wfnsym(n, (na, nb), ci[0], orbsym=orbsym) == aufbau
    other = aufbau ^ 1
    ref2, _ = _levels(h1, g, n, na, nb, sm, orbsym, other)
    e2, _ = sol.kernel(h1, g, n, (na, nb), orbsym=orbsym, wfnsym=other, davidson_only=True,
                       pspace_size=12)
    assert abs(e2 - ref2[0]) < 1e-9


def test_symm_impossible_irrep_raises():
    n, na, nb = 4, 2, 2
    orbsym = [0, 0, 1, 1]
    h1, g = _symm_problem(orbsym, 3)
    # every configuration's irrep is a XOR of ids in {0, 1}
    with pytest.raises(csf.ImpossibleSpinError):
        csf.FCISolver(smult=1, orbsym=orbsym, wfnsym=5).kernel(h1, g, n, (na, nb))
    with pytest.raises(ValueError):
        csf.CSFTransformer(n, na, nb, 1, orbsym=[0, 0, 1], wfnsym=0)
    with pytest.raises(RuntimeError):
        csf.CSFTransformer(n, na, nb, 1, orbsym=[0, 0, 9, 1], wfnsym=0)


def test_irrep_count_formula():
    """Independent count: the number of CSFs of irrep w is the sum over
    configurations of irrep w of the spin-coupling count of their open
    shells (binomial formula), with configurations enumerated here."""
    from math import comb
    n, na, nb, sm = 6, 3, 3, 3
    orbsym = ORBSYM6
    ne = na + nb
    s2 = sm - 1
    counts = numpy.zeros(8, dtype=int)
    for nd in range(ne // 2 + 1):
        no = ne - 2 * nd
        if no > n - nd or no < s2 or (no - s2) % 2:
            continue
        ncoup = comb(no, (no - s2) // 2) - comb(no, (no - s2) // 2 - 1) if (no - s2) // 2 > 0 \
            else 1
        for d in itertools.combinations(range(n), nd):
            rest = [p for p in range(n) if p not in d]
            for o in itertools.combinations(rest, no):
                w = 0
                for p in o:
                    w ^= orbsym[p]
                counts[w] += ncoup
    for w in range(8):
        t = csf.CSFTransformer(n, na, nb, sm, orbsym=orbsym, wfnsym=w)
        assert t.nsym == counts[w], (w, t.nsym, counts[w])
        t.close()


def test_pack_det_and_ncsf_under_symmetry():
    n, na, nb, sm, orbsym = 6, 3, 3, 1, ORBSYM6
    det_irr = _det_irreps(n, na, nb, orbsym)
    rng = numpy.random.default_rng(4)
    x = rng.standard_normal((2, det_irr.size))
    t = csf.CSFTransformer(n, na, nb, sm, orbsym=orbsym, wfnsym=2)
    assert t.ndeta == t.ndetb == 20 and t.ncsf == t.nsym == 40 and t.ncsf_full == 175
    p = t.pack_det(x)
    numpy.testing.assert_array_equal(p, x[:, det_irr == 2])
    numpy.testing.assert_array_equal(t.unpack_det(p), numpy.where(det_irr == 2, x, 0.0))
    pf = t.pack_det(numpy.asfortranarray(x.T), order="F")
    numpy.testing.assert_array_equal(pf, x[:, det_irr == 2].T)
    t.close()


@pytest.mark.parametrize("n,na,nb,sm", [(6, 3, 3, 1), (5, 3, 2, 2), (6, 4, 2, 3)])
def test_project_civec_is_the_spin_projector(n, na, nb, sm):
    S2 = brute.spin_square_matrix(n, (na, nb))
    w, v = numpy.linalg.eigh(0.5 * (S2 + S2.T))
    s = 0.5 * (sm - 1)
    P = v[:, abs(w - s * (s + 1)) < 1e-8]
    rng = numpy.random.default_rng(n + sm)
    x = rng.standard_normal(S2.shape[0])
    t = csf.CSFTransformer(n, na, nb, sm)
    y, norm = t.project_civec(x, normalize=False, return_norm=True)
    ref = P @ (P.T @ x)
    numpy.testing.assert_allclose(y, ref, rtol=0, atol=1e-12)
    assert abs(norm - numpy.linalg.norm(ref)) < 1e-12
    yn = t.project_civec(x)
    numpy.testing.assert_allclose(yn, ref / numpy.linalg.norm(ref), rtol=0, atol=1e-12)
    t.close()


@pytest.mark.parametrize("n,na,nb,sm", [(6, 3, 2, 2), (6, 3, 3, 3), (5, 3, 3, 1), (7, 4, 3, 4)])
def test_printable_csfstrings(n, na, nb, sm):
    """Each CSF's string against its determinant support (the '2's are
    the orbitals doubly occupied in every determinant, 'u'/'d' the singly
    occupied ones) and its coupling against the CSF oracle's genealogical
    strings, in order within each configuration (lowest SOMO first)."""
    import oracle
    import oracle.csf as ocsf
    t = csf.CSFTransformer(n, na, nb, sm)
    strs = t.printable_csfstring(list(range(t.ncsf)))
    assert t.printable_csfstring(0) == strs[0]
    assert strs == csf.printable_csfstring(n, na, nb, sm, numpy.arange(t.ncsf))
    sa, sb = oracle.make_strings(n, na), oracle.make_strings(n, nb)
    pos = 0
    while pos < t.ncsf:
        x = numpy.zeros(t.ncsf)
        x[pos] = 1.0
        d = t.vec_csf2det(x, normalize=False).reshape(len(sa), len(sb))
        ia, ib = numpy.nonzero(abs(d) > 1e-14)
        docc = {int(sa[i]) & int(sb[j]) for i, j in zip(ia, ib)}
        socc = {int(sa[i]) ^ int(sb[j]) for i, j in zip(ia, ib)}
        assert len(docc) == 1 and len(socc) == 1
        docc, socc = docc.pop(), socc.pop()
        nspin = bin(socc).count("1")
        coups = ocsf.coupling_strings(nspin, sm)
        for k, cp in enumerate(coups):
            s = strs[pos + k]
            somo = [p for p in range(n) if (socc >> p) & 1]
            for p in range(n):
                c = s[n - 1 - p]
                if (docc >> p) & 1:
                    assert c == "2"
                elif (socc >> p) & 1:
                    bit = (int(cp) >> somo.index(p)) & 1
                    assert c == ("u" if bit else "d"), (s, bin(int(cp)))
                else:
                    assert c == "0"
        pos += len(coups)
    # the largest components come first
    rng = numpy.random.default_rng(1)
    v = rng.standard_normal(t.ncsf)
    ps, cs = t.printable_largest_csf(v, 3)
    order = numpy.argsort(-abs(v))[:3]
    numpy.testing.assert_array_equal(cs[0], v[order])
    assert list(ps[0]) == [strs[i] for i in order]
    t.close()


def test_module_level_csf_api():
    """csf.get_init_guess / pspace / kernel (csf.py:51-64, 246, 368) and
    csf_symm.FCISolver as module functions over a solver."""
    n, na, nb, sm, orbsym = 6, 3, 3, 1, ORBSYM6
    h1, g = _symm_problem(orbsym, 21)
    sol = pkg.csf_symm.FCISolver(smult=sm, orbsym=orbsym, wfnsym=0)
    t = sol.check_transformer_cache(n, (na, nb))
    hd = sol.make_hdiag_csf(h1, g, n, (na, nb))
    guesses = csf.get_init_guess(n, (na, nb), 2, hd, t)
    assert len(guesses) == 2
    det_irr = _det_irreps(n, na, nb, orbsym)
    for v in guesses:
        assert abs(v.reshape(-1)[det_irr != 0]).max(initial=0.0) == 0.0
    addr, h0 = csf.pspace(sol, h1, g, n, (na, nb), t, npsp=10)
    assert h0.shape == (addr.size, addr.size) and addr.size <= 10
    e, _ = csf.kernel(sol, h1, g, n, (na, nb), smult=sm, nroots=1)
    ref, _ = _levels(h1, g, n, na, nb, sm, orbsym, 0)
    assert abs(e - ref[0]) < 1e-9
    with pytest.raises(NotImplementedError):
        csf.kernel(sol, h1, g, n, (na, nb), idx_sym=numpy.ones(t.ncsf_full, dtype=bool))
    lines = []
    sol.transformer.print_config(lines.append)
    assert any("ncsf = %d" % t.ncsf in x for x in lines)


def test_module_level_transforms():
    """transform_civec_det2csf / _csf2det, check_spinstate_norm and
    project_civec_csf (csfstring.py:297-487) against the transformer and
    the brute-force spin projector."""
    n, na, nb, sm = 6, 3, 3, 3
    rng = numpy.random.default_rng(8)
    S2 = brute.spin_square_matrix(n, (na, nb))
    w, v = numpy.linalg.eigh(0.5 * (S2 + S2.T))
    P = v[:, abs(w - 2.0) < 1e-8]
    x = rng.standard_normal(S2.shape[0])
    ref = P @ (P.T @ x)
    y, norm = csf.transform_civec_det2csf(x, n, na, nb, sm, do_normalize=False)
    assert abs(norm - numpy.linalg.norm(ref)) < 1e-12
    assert abs(csf.check_spinstate_norm(x, n, na, nb, sm) - norm) < 1e-12
    d, _ = csf.transform_civec_csf2det(y, n, na, nb, sm, do_normalize=False)
    numpy.testing.assert_allclose(d, ref, rtol=0, atol=1e-12)
    p, pn = csf.project_civec_csf(x.reshape(20, 20), n, na, nb, sm)
    assert p.shape == (20, 20) and abs(pn - numpy.linalg.norm(ref)) < 1e-12
    numpy.testing.assert_allclose(p.reshape(-1), ref / numpy.linalg.norm(ref), rtol=0, atol=1e-12)


@pytest.mark.parametrize("sym", [False, True])
def test_supports_match_unit_transforms(sym):
    """fcis_csf_support rows equal the nonzero entries of csf2det of unit
    CSF vectors (packed indices under symmetry)."""
    import torch
    n, na, nb, sm = 7, 4, 3, 2
    t = csf.CSFTransformer(n, na, nb, sm, orbsym=ORBSYM7 if sym else None,
                           wfnsym=3 if sym else None)
    idx = torch.tensor([0, 1, 5, t.ncsf // 2, t.ncsf - 1], device="cuda")
    A, C, cnt = t.supports(idx)
    for row, i in enumerate(idx.tolist()):
        x = numpy.zeros(t.ncsf)
        x[i] = 1.0
        d = t.vec_csf2det(x, normalize=False)
        nz = numpy.nonzero(d)[0]
        k = int(cnt[row])
        got_a = A[row, :k].cpu().numpy()
        got_c = C[row, :k].cpu().numpy()
        order = numpy.argsort(got_a)
        numpy.testing.assert_array_equal(got_a[order], nz)
        numpy.testing.assert_allclose(got_c[order], d[nz], rtol=0, atol=1e-15)
    t.close()


def test_csf_symm_solver_irrep_names_and_restored_attributes():
    """csf_symm.FCISolver: an irrep name with the molecule's group gives the
    irrep id's spectrum; kernel kwargs apply to that call only; the guess
    needs orbsym and wfnsym; csf_solver(symm=True) returns this solver."""
    from types import SimpleNamespace
    cs = pkg.csf_symm
    n, na, nb, sm = 6, 3, 3, 1
    h1, g = _symm_problem(ORBSYM6, 11)
    mol = SimpleNamespace(groupname="C2v", symmetry=True)
    sol = pkg.csf_fci.csf_solver(mol, smult=sm)
    assert isinstance(sol, cs.FCISolver)
    for name in ("A1", "B2"):
        w = cs.irrep_name2id("C2v", name)
        want, _ = _levels(h1, g, n, na, nb, sm, ORBSYM6, w)
        e, ci = sol.kernel(h1, g, n, (na, nb), orbsym=ORBSYM6, wfnsym=name, nroots=2)
        numpy.testing.assert_allclose(e, want[:2], atol=1e-9)
        assert sol.orbsym is None and sol.wfnsym is None
        irr = _det_irreps(n, na, nb, ORBSYM6)
        assert abs(ci[0].reshape(-1)[irr != w]).max() < 1e-12
    with pytest.raises(AssertionError):
        sol.get_init_guess(n, (na, nb), 1, numpy.zeros(10))
    with pytest.raises(NotImplementedError):
        sol.kernel(h1, g, n, (na, nb), orbsym=ORBSYM6, wfnsym=12)

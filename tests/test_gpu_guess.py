"""Initial guess and diagonal preconditioner of the solvers against the
oracle's restatement of the reference's (pyscf/mcscf/fci_gas.py:46-88,
:881-902, :1337-1342), itself pinned by the SplitMix64 known answers and
numpy.argpartition (tests/test_oracle.py).

Guesses: the same address per root in the reference's argpartition order
and bit-identical noise, for nroots 1..4 on diagonals with distinct values;
on a real CAS diagonal with tied (alpha <-> beta mirror) entries, the same
set of lowest values, ordered by (value, address).  Preconditioner: equal to
dx / (hdiag - e) with every |hdiag - e| < level_shift replaced by
+level_shift, bit for bit, through fcis_precond and fcis_precond_list."""

import numpy
import pytest

import oracle
from conftest import has_gpu, package

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not has_gpu(), reason="needs a HIP device")]

pkg = package()


@pytest.mark.parametrize("n", [7, 50, 1000, 300000])
@pytest.mark.parametrize("nroots", [1, 2, 3, 4])
def test_init_guess_matches_reference_restatement(n, nroots):
    hd = numpy.random.default_rng(n + nroots).normal(size=n)
    want_addr, want = oracle.init_guess(hd, nroots)
    got = pkg.FCISolver().get_init_guess(None, None, nroots, hd)
    assert len(got) == len(want) == min(nroots, n)
    for a, g, w in zip(want_addr, got, want):
        assert int(numpy.argmax(g)) == int(a)
        numpy.testing.assert_array_equal(g, w)


def test_init_guess_whole_range_when_space_is_small():
    hd = numpy.array([2.0, -1.0, 0.5])
    want_addr, want = oracle.init_guess(hd, 3)
    assert list(want_addr) == [0, 1, 2]
    got = pkg.FCISolver().get_init_guess(None, None, 3, hd)
    for g, w in zip(got, want):
        numpy.testing.assert_array_equal(g, w)


def test_init_guess_ties_on_a_cas_diagonal():
    """na == nb: hdiag[Ia, Ib] == hdiag[Ib, Ia] exactly.  The lowest values
    agree with argpartition's; ties are ordered by address."""
    n, ne = 8, (4, 4)
    h1e, eri = pkg.synthetic.make_problem(n, ne, seed=1, with_ci=False)
    hd = pkg.make_hdiag(h1e, eri, n, ne).reshape(-1)
    k = 6
    got = pkg.FCISolver().get_init_guess(n, ne, k, hd)
    addrs = [int(numpy.argmax(g)) for g in got]
    ref_addr, _ = oracle.init_guess(hd, k)
    numpy.testing.assert_array_equal(numpy.sort(hd[addrs]), numpy.sort(hd[ref_addr]))
    key = [(hd[a], a) for a in addrs]
    assert key == sorted(key)
    _, want = oracle.init_guess(hd, k, addrs=addrs)
    for g, w in zip(got, want):
        numpy.testing.assert_array_equal(g, w)


def test_gas_and_batch_guesses_use_the_same_rule():
    from conftest import GOLDEN
    import json
    import os
    with numpy.load(os.path.join(GOLDEN, "gas", "gas_ras_n10.npz")) as f:
        d = {k: f[k] for k in f.files}
    orbs = tuple(int(x) for x in d["gas_orbs"])
    ne = tuple(int(x) for x in d["nelec"])
    sol = pkg.fci_gas.FCISolver(gas_orbs=orbs, gas_restr=json.loads(str(d["restr"])),
                                gas_restr_type=str(d["restr_type"]))
    hd = d["hdiag"].reshape(-1) + 1e-9 * numpy.arange(d["hdiag"].size)   # distinct values
    want_addr, want = oracle.init_guess(hd, 3)
    got = sol.get_init_guess(sum(orbs), ne, 3, hd)
    for g, w in zip(got, want):
        numpy.testing.assert_array_equal(g, w)


def test_precond_matches_reference_rule():
    import torch
    n = 4096
    rng = numpy.random.default_rng(3)
    hd = rng.normal(size=n)
    e, ls = 0.25, 1e-3
    # denominators just inside / outside the level shift, of both signs
    hd[:6] = e + numpy.array([5e-4, -5e-4, 1e-3, -1e-3, 2e-3, -2e-3])
    hd[6] = e
    dx = rng.normal(size=n)
    want = oracle.precond(dx, hd, e, ls)
    hdd = torch.from_numpy(hd).cuda()
    dxd = torch.from_numpy(dx).cuda()
    out = torch.empty_like(dxd)
    lib = pkg._lib.load()
    pkg._lib.check(lib.fcis_precond(n, pkg._lib.vptr(hdd), e, ls, pkg._lib.vptr(dxd),
                                    pkg._lib.vptr(out), pkg._lib.stream_ptr()), "fcis_precond")
    numpy.testing.assert_array_equal(out.cpu().numpy(), want)
    out2 = torch.empty_like(dxd)
    pkg.batch.BlockOps().precond([(hdd.data_ptr(), dxd.data_ptr(), out2.data_ptr(), n, e)], ls)
    numpy.testing.assert_array_equal(out2.cpu().numpy(), want)


def test_lowest_address_single_root_ties():
    """One root: the lowest value's lowest address (the rule the counts above
    one follow), also when the minimum is tied."""
    import torch
    rng = numpy.random.default_rng(3)
    hd = rng.normal(size=5000)
    hd[3700] = hd[1300] = hd.min() - 1.0
    d = torch.from_numpy(hd).cuda()
    assert pkg.direct_spin1.lowest_addresses(d, 1).tolist() == [1300]
    assert pkg.direct_spin1.lowest_addresses(d, 2).tolist() == [1300, 3700]
    hd[1300] += 0.5
    d = torch.from_numpy(hd).cuda()
    assert pkg.direct_spin1.lowest_addresses(d, 1).tolist() == [3700]

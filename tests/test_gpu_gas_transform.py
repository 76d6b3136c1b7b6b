"""GPU parity of transform_ci_within_gas (pyscf/mcscf/fci_gas.py:1003-1091)
on the device (fcis_rotate_fibers).  Pinned by what the rotation must do,
not by its own formula: for mo_new = mo_old @ u the transformed vector's
RDMs are the rotated RDMs, dm1' = u^T dm1 u (and the same on all four
indices of dm2), with the RDMs from the oracle (oracle/gas.py, pinned to
libfci_gas); the norm is kept; u then u^T is the identity; the energy in
the rotated integrals is unchanged.  Tolerances 1e-12 (vectors, RDMs),
1e-10 (energies)."""

import json
import os

import numpy
import pytest

from oracle import gas as og
from conftest import GOLDEN, has_gpu, package

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not has_gpu(), reason="needs a HIP device")]

pkg = package()
fg = pkg.fci_gas


def load(name):
    with numpy.load(os.path.join(GOLDEN, "gas", name + ".npz")) as f:
        d = {k: f[k] for k in f.files}
    d["gas_orbs"] = tuple(int(x) for x in d["gas_orbs"])
    d["nelec"] = tuple(int(x) for x in d["nelec"])
    d["restr_type"] = str(d["restr_type"])
    d["restr"] = json.loads(str(d["restr"]))
    return d


def _block_rotation(orbs, seed, reflect=True, permute=True):
    rng = numpy.random.default_rng(seed)
    n = sum(orbs)
    u = numpy.zeros((n, n))
    o = 0
    for g, k in enumerate(orbs):
        q, _ = numpy.linalg.qr(rng.standard_normal((k, k)))
        if reflect and g == 0:
            q[:, 0] *= -1
        if permute and g == len(orbs) - 1 and k > 1:
            q = q[:, ::-1]
        u[o:o + k, o:o + k] = q
        o += k
    return u


def _rot4(t, u):
    return numpy.einsum("ijkl,ip,jq,kr,ls->pqrs", t, u, u, u, u, optimize=True)


@pytest.mark.parametrize("name", ["gas_sg_n4", "gas_ras_n4", "gas_cum_n3", "gas_sg_n8",
                                  "gas_ras_n10", "gas_cum4_n10"])
def test_rdms_rotate(name):
    d = load(name)
    n = sum(d["gas_orbs"])
    L = og.GasLayout(d["gas_orbs"], d["nelec"], d["blocks"])
    sol = fg.FCISolver(gas_orbs=d["gas_orbs"], gas_restr=d["restr"],
                       gas_restr_type=d["restr_type"])
    u = _block_rotation(d["gas_orbs"], len(name))
    c = d["ci"] / numpy.linalg.norm(d["ci"])
    c1 = sol.transform_ci_within_gas(c, n, d["nelec"], u)
    assert c1.shape == c.shape
    assert abs(numpy.linalg.norm(c1) - 1.0) < 1e-12
    (a0, b0), (aa0, ab0, bb0) = og.gas_trans_rdm12s(c, c, L)
    (a1, b1), (aa1, ab1, bb1) = og.gas_trans_rdm12s(c1, c1, L)
    numpy.testing.assert_allclose(a1, u.T @ a0 @ u, rtol=0, atol=1e-12)
    numpy.testing.assert_allclose(b1, u.T @ b0 @ u, rtol=0, atol=1e-12)
    for x1, x0 in ((aa1, aa0), (ab1, ab0), (bb1, bb0)):
        numpy.testing.assert_allclose(x1, _rot4(x0, u), rtol=0, atol=1e-12)
    back = sol.transform_ci_within_gas(c1, n, d["nelec"], u.T)
    numpy.testing.assert_allclose(back, c, rtol=0, atol=1e-12)
    numpy.testing.assert_array_equal(sol.transform_ci_within_gas(c, n, d["nelec"], numpy.eye(n)), c)


def test_energy_invariant_in_a_large_cas():
    """One subspace of 12 orbitals (853,776 determinants): E(c; h, g) =
    E(c'; u^T h u, g rotated)."""
    n, ne = 12, (6, 6)
    h1e, eri, c = pkg.synthetic.make_problem(n, ne, seed=2)
    import oracle
    g = oracle.unpack_pair_matrix(oracle.pair_matrix(eri, n), n)
    sol = fg.FCISolver(gas_orbs=(n,))
    u = _block_rotation((n,), 9)
    c = c.reshape(-1) / numpy.linalg.norm(c)
    c1 = sol.transform_ci_within_gas(c, n, ne, u)
    e0 = sol.energy(h1e, g, c, n, ne)
    e1 = sol.energy(u.T @ h1e @ u, _rot4(g, u), c1, n, ne)
    assert abs(e1 - e0) < 1e-10 * max(1.0, abs(e0))


def test_rejections():
    d = load("gas_ras_n4")
    n = sum(d["gas_orbs"])
    sol = fg.FCISolver(gas_orbs=d["gas_orbs"], gas_restr=d["restr"],
                       gas_restr_type=d["restr_type"])
    u = numpy.eye(n)
    u[[0, 1]] = u[[1, 0]]                     # mixes subspaces 0 and 1
    with pytest.raises(ValueError):
        sol.transform_ci_within_gas(d["ci"], n, d["nelec"], u)
    with pytest.raises(ValueError):
        sol.transform_ci_within_gas(d["ci"], n, d["nelec"], 2 * numpy.eye(n))
    with pytest.raises(NotImplementedError):
        sol.transform_ci_for_orbital_rotation(d["ci"], n, d["nelec"], numpy.eye(n))

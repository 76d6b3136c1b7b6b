"""CI side of GASSCF's orbital/CI Newton coupling on the GPU (the reference's
pyscf/mcscf/gasscf.py:59-97, 213-600).

GASSCF drives pyscf's joint orbital/CI Newton optimiser
(``newton_casscf.gen_g_hop``, pyscf core) through fcisolver calls:
``contract_2e`` on the current and trial CI vectors, ``make_rdm12`` /
``trans_rdm12`` for the orbital blocks, and, with a spin penalty, ``S^2``
contractions.  This module holds the reference's GASSCF-owned pieces of
that coupling:

* ``gen_g_hop``: wraps the Newton gradient / update / Hessian-vector /
  diagonal of the base driver and adds the spin-penalty CI terms
  (gasscf.py:213-289), with every S^2 application on the device
  (GasSpinPlan.penalty);
* ``_GASFCISolver``: the solver shell with Newton-owned plans (contraction
  plans keyed by a digest of the absorbed integrals, LRU of three; the RDM
  and S^2 plans), single-root unwrapping of ``[ci]`` lists and the hooks
  native Newton probes (gasscf.py:292-538);
* ``adapt_solver`` / ``new_gas_solver`` (gasscf.py:541-600) and the
  symmetry check (gasscf.py:63-79).

The orbital blocks of the Newton step (integral transforms, the orbital
Hessian) belong to pyscf core and are not rebuilt here: ``gen_g_hop`` takes
the base driver from ``pyscf.mcscf.newton_casscf`` or from its ``base``
argument.
"""

import copy as _copy
import hashlib
from collections import OrderedDict

import numpy

from . import addons_gas
from . import direct_spin1 as _ds
from . import fci_gas


def _unsupported(feature):
    raise NotImplementedError(feature + " is not implemented for GASSCF")


def _check_symmetry(obj, mo_coeff=None, mol=None):
    """Reject point-group symmetry (gasscf.py:63-79): the GAS sectors carry
    no irrep labels."""
    if mol is not None and getattr(mol, "symmetry", False):
        _unsupported("molecular point-group symmetry")
    for source in (obj, getattr(obj, "_scf", None), getattr(obj, "fcisolver", None)):
        if source is None:
            continue
        if getattr(getattr(source, "mol", None), "symmetry", False):
            _unsupported("molecular point-group symmetry")
        for name in ("extrasym", "orbsym", "wfnsym"):
            if getattr(source, name, None) is not None:
                _unsupported(name + " symmetry constraints")
        if getattr(getattr(source, "mo_coeff", None), "orbsym", None) is not None:
            _unsupported("orbital symmetry labels (orbsym)")
    if getattr(mo_coeff, "orbsym", None) is not None:
        _unsupported("orbital symmetry labels (orbsym)")


def _effective_nelecas(mc):
    f = getattr(mc, "_effective_nelecas", None)
    if f is not None:
        return f()
    return _ds._unpack_nelec(mc.nelecas, getattr(mc.fcisolver, "spin", None))


def _base_gen_g_hop():
    try:
        from pyscf.mcscf import newton_casscf
    except ImportError as err:
        raise ImportError("gen_g_hop needs pyscf's newton_casscf.gen_g_hop (pyscf core) or "
                          "an explicit base= driver") from err
    return newton_casscf.gen_g_hop


def gen_g_hop(mc, mo, ci0, eris, verbose=None, base=None):
    """Newton (gradient, update, hop, hdiag) with the spin penalty's CI
    terms (gasscf.py:213-289).  For each root c of weight w, p = <c|P|c>
    and r = P c - p c with P the penalty operator:

        gradient[c]  += 2 w r
        hdiag[c]     += 2 w (diag(P) - p - 2 r c)
        hop(v)[c]    += 2 w [(P - p) v - r (c.v) - c (r.v)]
        update(u,ci) += 2 w (P c' - (c'.P c') c'),  c' = ci / |ci|

    The orbital rows are the base driver's.  ``base`` has the signature of
    ``newton_casscf.gen_g_hop``."""
    _check_symmetry(mc, mo)
    base = base if base is not None else _base_gen_g_hop()
    gradient, update, hop, hdiag = base(mc, mo, ci0, eris, verbose)
    solver = mc.fcisolver
    nelecas = _effective_nelecas(mc)
    parameters = fci_gas._spin_penalty_parameters(solver, mc.ncas, nelecas)
    if parameters is None or parameters[0] == 0:
        return gradient, update, hop, hdiag
    if getattr(solver, "cache_plans", False) and hasattr(solver, "_get_spin_plan"):
        plan = solver._get_spin_plan(mc.ncas, mc.nelecas)
    else:
        plan = solver.make_spin_plan(mc.ncas, mc.nelecas)

    def penalty(vector):
        return plan.penalty(vector, parameters)

    nroots = int(getattr(solver, "nroots", 1))
    roots = [ci0] if nroots == 1 else ci0
    roots = [numpy.asarray(c, dtype=numpy.float64).ravel() for c in roots]
    weights = tuple(getattr(mc, "weights", (1.0,)))
    ngorb = gradient.size - sum(c.size for c in roots)
    diagonal = fci_gas._spin_penalty_diagonal(plan.diagonal_vector(), parameters)
    gradient = gradient.copy()
    hdiag = hdiag.copy()
    terms = []
    start = ngorb
    for c, weight in zip(roots, weights):
        block = slice(start, start + c.size)
        start += c.size
        if weight == 0:
            continue
        pc = penalty(c)
        energy = float(c.dot(pc))
        residual = pc - energy * c
        gradient[block] += 2 * weight * residual
        hdiag[block] += 2 * weight * (diagonal - energy - 2 * residual * c)
        terms.append((block, c, weight, energy, residual))

    def penalized_hop(vector):
        result = hop(vector)
        for block, c, weight, energy, residual in terms:
            v = vector[block]
            result[block] += 2 * weight * (penalty(v) - energy * v - residual * c.dot(v)
                                           - c * residual.dot(v))
        return result

    def penalized_update(u, ci):
        result = update(u, ci)
        current = [ci] if nroots == 1 else ci
        start = ngorb
        for c, weight in zip(current, weights):
            c = numpy.asarray(c, dtype=numpy.float64).ravel()
            block = slice(start, start + c.size)
            start += c.size
            if weight == 0:
                continue
            c = c / numpy.linalg.norm(c)
            pc = penalty(c)
            result[block] += 2 * weight * (pc - c.dot(pc) * c)
        return result

    return gradient, penalized_update, penalized_hop, hdiag


class _GASFCISolver(fci_gas.FCISolver):
    """GAS solver shell with Newton-owned plans (gasscf.py:292-538).

    Contraction plans are cached per absorbed-integral digest (at most three,
    least recently used evicted), the RDM and S^2 plans per GAS topology;
    a topology change closes them all.  Copies start with empty caches.  An
    explicit ``plan=`` is validated and borrowed for that call only."""

    _MAX_CONTRACT_PLANS = 3

    # native Newton probes these with getattr: no CAS link strings, and an
    # orbital rotation does not preserve a restricted GAS space
    gen_linkstr = None
    transform_ci_for_orbital_rotation = None

    def __init__(self, mol=None, gas_orbs=None, gas_restr=None,
                 gas_restr_type=addons_gas.GAS_RESTR_SPIN_SUPERGROUP, lib=None, *,
                 cache_plans=True):
        super().__init__(mol, gas_orbs=gas_orbs, gas_restr=gas_restr,
                         gas_restr_type=gas_restr_type, lib=lib)
        self.cache_plans = bool(cache_plans)
        self._init_plan_cache()

    def _init_plan_cache(self):
        """Detach (without closing) any cached plans: used right after a
        shallow copy, whose attributes may still belong to the source."""
        self._topology_key = None
        self._contract_space = None
        self._contract_plans = OrderedDict()
        self._rdm_plan_cache = None
        self._spin_plan = None

    def _space_key(self, norb, nelec):
        orbs, nelec, blocks = self._space_spec(norb, nelec)
        return (tuple(int(v) for v in orbs), tuple(int(v) for v in nelec),
                tuple(tuple(int(x) for x in row) for row in numpy.asarray(blocks)),
                id(self.lib))

    def _ensure_topology(self, norb, nelec):
        key = self._space_key(norb, nelec)
        if getattr(self, "_topology_key", None) != key:
            self.close()
            self._topology_key = key
        return key

    @staticmethod
    def _contract_eri_key(eri):
        arr = numpy.ascontiguousarray(eri, dtype=numpy.float64)
        digest = hashlib.sha256(arr.view(numpy.uint8)).digest()
        return arr, (tuple(int(v) for v in arr.shape), digest)

    def _get_contract_plan(self, eri, norb, nelec):
        """Newton-owned contraction plan for one absorbed Hamiltonian."""
        eri, key = self._contract_eri_key(eri)
        self._ensure_topology(norb, nelec)
        if self._contract_space is None:
            self._contract_space = self.make_space(norb, nelec, compress_links=True)
        plan = self._contract_plans.pop(key, None)
        if plan is None:
            if len(self._contract_plans) >= self._MAX_CONTRACT_PLANS:
                _, evicted = self._contract_plans.popitem(last=False)
                evicted.close()
            plan = fci_gas.GasContractPlan(self._contract_space, eri.copy())
        self._contract_plans[key] = plan
        return plan

    def _get_rdm_plan(self, norb, nelec):
        self._ensure_topology(norb, nelec)
        if self._rdm_plan_cache is None:
            self._rdm_plan_cache = self.make_rdm_plan(norb, nelec)
        return self._rdm_plan_cache

    def _rdm_plan(self, norb, nelec, plan):
        if plan is None and self.cache_plans:
            plan = self._get_rdm_plan(norb, nelec)
        return super()._rdm_plan(norb, nelec, plan)

    def _get_spin_plan(self, norb, nelec):
        self._ensure_topology(norb, nelec)
        if self._spin_plan is None:
            self._spin_plan = self.make_spin_plan(norb, nelec)
        return self._spin_plan

    def close(self):
        """Release the cached plans; repeated calls are safe."""
        plans = getattr(self, "_contract_plans", None)
        if plans is not None:
            for p in list(plans.values()):
                p.close()
            plans.clear()
        space = getattr(self, "_contract_space", None)
        if space is not None:
            space.close()
            self._contract_space = None
        rdm = getattr(self, "_rdm_plan_cache", None)
        if rdm is not None:
            rdm.close()
            self._rdm_plan_cache = None
        self._spin_plan = None
        self._topology_key = None

    def contract_2e(self, eri, fcivec, norb, nelec, link_index=None, *args, **kwargs):
        """P H P c, reusing a Newton-owned plan for repeated integrals."""
        plan = kwargs.pop("plan", None)
        if plan is not None:
            return super().contract_2e(eri, fcivec, norb, nelec, link_index, *args, plan=plan,
                                       **kwargs)
        compress_links = bool(kwargs.pop("compress_links", True))
        if args or kwargs or not self.cache_plans or not compress_links:
            return super().contract_2e(eri, fcivec, norb, nelec, link_index, *args, **kwargs)
        plan = self._get_contract_plan(eri, norb, nelec)
        return super().contract_2e(eri, fcivec, norb, nelec, link_index, plan=plan)

    @staticmethod
    def _as_state_specific_ci(ci):
        """Unwrap native Newton's ``[ci]`` form of one state."""
        if isinstance(ci, (list, tuple)):
            if len(ci) != 1:
                _unsupported("multiroot CI density dispatch")
            return ci[0]
        return ci

    def spin_square(self, ci, norb, nelec, *args, **kwargs):
        ci = self._as_state_specific_ci(ci)
        return super().spin_square(ci, norb, nelec)

    def make_rdm1s(self, ci, norb, nelec, link_index=None, *, plan=None):
        return super().make_rdm1s(self._as_state_specific_ci(ci), norb, nelec, link_index,
                                  plan=plan)

    def make_rdm1(self, ci, norb, nelec, link_index=None, *, plan=None):
        return super().make_rdm1(self._as_state_specific_ci(ci), norb, nelec, link_index,
                                 plan=plan)

    def make_rdm12s(self, ci, norb, nelec, link_index=None, reorder=True, *, plan=None):
        return super().make_rdm12s(self._as_state_specific_ci(ci), norb, nelec, link_index,
                                   reorder, plan=plan)

    def make_rdm12(self, ci, norb, nelec, link_index=None, reorder=True, *, plan=None):
        return super().make_rdm12(self._as_state_specific_ci(ci), norb, nelec, link_index,
                                  reorder, plan=plan)

    def trans_rdm1s(self, cibra, ciket, norb, nelec, link_index=None, *, plan=None):
        return super().trans_rdm1s(self._as_state_specific_ci(cibra),
                                   self._as_state_specific_ci(ciket), norb, nelec, link_index,
                                   plan=plan)

    def trans_rdm1(self, cibra, ciket, norb, nelec, link_index=None, *, plan=None):
        return super().trans_rdm1(self._as_state_specific_ci(cibra),
                                  self._as_state_specific_ci(ciket), norb, nelec, link_index,
                                  plan=plan)

    def trans_rdm12s(self, cibra, ciket, norb, nelec, link_index=None, reorder=True, *,
                     plan=None):
        return super().trans_rdm12s(self._as_state_specific_ci(cibra),
                                    self._as_state_specific_ci(ciket), norb, nelec, link_index,
                                    reorder, plan=plan)

    def trans_rdm12(self, cibra, ciket, norb, nelec, link_index=None, reorder=True, *,
                    plan=None):
        return super().trans_rdm12(self._as_state_specific_ci(cibra),
                                   self._as_state_specific_ci(ciket), norb, nelec, link_index,
                                   reorder, plan=plan)

    def contract_ss(self, fcivec, norb, nelec):
        """S^2 c, reusing the Newton-owned S^2 plan."""
        fcivec = self._as_state_specific_ci(fcivec)
        if not self.cache_plans:
            return super().contract_ss(fcivec, norb, nelec)
        return self._get_spin_plan(norb, nelec).contract(fcivec)

    def copy(self):
        result = _copy.copy(self)
        result._init_plan_cache()
        return result


def _validated_gas_orbs(gas_orbs):
    orbs = tuple(int(x) for x in numpy.asarray(gas_orbs).reshape(-1))
    if not orbs or any(x <= 0 for x in orbs) or any(
            float(a) != float(b) for a, b in zip(orbs, numpy.asarray(gas_orbs).reshape(-1))):
        raise ValueError("gas_orbs must be positive integers")
    return orbs


def new_gas_solver(mf, gas_orbs, gas_restr=None, gas_restr_type=None, cache_plans=None):
    """Solver for a GASSCF object built from a mean field (gasscf.py:547-558)."""
    if gas_orbs is None:
        raise ValueError("gas_orbs is required without an explicit GASCI solver")
    if gas_restr_type is None:
        gas_restr_type = addons_gas.GAS_RESTR_SPIN_SUPERGROUP
    return _GASFCISolver(getattr(mf, "mol", None), gas_orbs=_validated_gas_orbs(gas_orbs),
                         gas_restr=gas_restr, gas_restr_type=gas_restr_type,
                         cache_plans=True if cache_plans is None else bool(cache_plans))


def adapt_solver(fcisolver, cache_plans=None):
    """A Newton-owned shell for an explicit GAS solver (gasscf.py:572-600):
    settings copied, plans never shared with the input."""
    if isinstance(fcisolver, _GASFCISolver):
        solver = fcisolver.copy()
    elif type(fcisolver) is fci_gas.FCISolver:
        solver = _copy.copy(fcisolver)
        solver.__class__ = _GASFCISolver
        solver._init_plan_cache()
    else:
        _unsupported("external/non-GASCI solver adaptation")
    if solver.gas_orbs is None:
        raise ValueError("explicit GASCI solver must define gas_orbs")
    solver.gas_orbs = _validated_gas_orbs(solver.gas_orbs)
    if solver.gas_restr_type is None:
        solver.gas_restr_type = addons_gas.GAS_RESTR_SPIN_SUPERGROUP
    if cache_plans is None:
        solver.cache_plans = bool(getattr(fcisolver, "cache_plans", True))
    else:
        solver.cache_plans = bool(cache_plans)
    return solver

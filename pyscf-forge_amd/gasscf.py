"""CI side of the GASSCF orbital/CI Newton coupling, on the GPU.

The reference's GASSCF (pyscf/mcscf/gasscf.py) drives pyscf's joint
orbital/CI Newton optimiser (``newton_casscf.gen_g_hop``, pyscf core)
through fcisolver calls: ``contract_2e`` on current and trial CI vectors,
``make_rdm12`` / ``trans_rdm12`` for the orbital blocks and, with a spin
penalty, ``S^2`` contractions.  This module holds the GASSCF-owned parts
of that coupling:

* ``gen_g_hop`` (the reference's gasscf.py:213-289): the Newton gradient,
  update, Hessian-vector product and diagonal of a base driver plus the CI
  terms of the spin penalty, every S^2 application on the device;
* ``_GASFCISolver`` (gasscf.py:292-538): a GAS solver that keeps the plans
  a Newton run reuses (contraction plans per absorbed Hamiltonian, RDM and
  S^2 plans per space) and accepts Newton's one-element ``[ci]`` lists;
* ``new_gas_solver`` / ``adapt_solver`` (gasscf.py:541-600) and the
  point-group rejection (gasscf.py:63-79).

The orbital blocks of the Newton step (integral transforms, orbital
Hessian) belong to pyscf core; ``gen_g_hop`` takes its base driver from
``pyscf.mcscf.newton_casscf`` or from the ``base`` argument.
"""

import copy as _copy
import hashlib

import numpy

from . import addons_gas
from . import direct_spin1 as _ds
from . import fci_gas


def _unsupported(feature):
    raise NotImplementedError(feature + " is not implemented for GASSCF")


# Point-group markers, refused rather than silently ignored (GAS sectors
# carry no irrep labels): (attribute path, refused when truthy rather than
# when merely present, what it means).
_SYMMETRY_MARKERS = (
    (("mol", "symmetry"), True, "molecular point-group symmetry"),
    (("extrasym",), False, "extrasym symmetry constraints"),
    (("orbsym",), False, "orbsym symmetry constraints"),
    (("wfnsym",), False, "wfnsym symmetry constraints"),
    (("mo_coeff", "orbsym"), False, "orbital symmetry labels (orbsym)"),
)


def _lookup(obj, path):
    for name in path:
        obj = getattr(obj, name, None)
        if obj is None:
            break
    return obj


def _check_symmetry(obj, mo_coeff=None, mol=None):
    """Refuse point-group symmetry on the GASSCF object, its mean field, its
    solver, the orbitals or the molecule (gasscf.py:63-79)."""
    owners = [o for o in (obj, getattr(obj, "_scf", None), getattr(obj, "fcisolver", None))
              if o is not None]
    if mol is not None and getattr(mol, "symmetry", False):
        _unsupported("molecular point-group symmetry")
    for owner in owners:
        for path, truthy, what in _SYMMETRY_MARKERS:
            flag = _lookup(owner, path)
            if (bool(flag) if truthy else flag is not None):
                _unsupported(what)
    if _lookup(mo_coeff, ("orbsym",)) is not None:
        _unsupported("orbital symmetry labels (orbsym)")


def _active_electrons(mc):
    own = getattr(mc, "_effective_nelecas", None)
    if own is not None:
        return own()
    return _ds._unpack_nelec(mc.nelecas, getattr(mc.fcisolver, "spin", None))


def _newton_base():
    try:
        from pyscf.mcscf import newton_casscf
    except ImportError as err:
        raise ImportError("gen_g_hop needs pyscf's newton_casscf.gen_g_hop (pyscf core) or "
                          "an explicit base= driver") from err
    return newton_casscf.gen_g_hop


class _PenaltyNewtonTerms:
    """CI-block contributions of the spin penalty to one Newton step.

    Root k with weight w_k adds w_k f(c_k), f(c) = <c|P|c> / <c|c>.  At a
    normalised c, with p = f(c) and r = P c - p c (the penalty residual):

        d f / dc          = 2 r
        d^2 f / dc dc     = 2 (P - p) - 4 (c r^T + r c^T).

    The Newton driver's CI blocks use the symmetric form
    2 [(P - p) v - r (c.v) - c (r.v)] for the Hessian-vector product, which
    equals the exact second derivative on the tangent space of the unit
    sphere (v orthogonal to c, result projected off c), and its diagonal
    2 (diag P - p - 2 r c) entrywise; the update at a new CI vector is the
    gradient at its normalised direction.  P is applied on the device
    (GasSpinPlan.penalty); diag P is SpinPenalty.diagonal."""

    def __init__(self, plan, penalty, roots, weights, offset):
        self.plan, self.penalty = plan, penalty
        self.offset = offset
        self.active = []                  # (slice, weight, c, p, r) of the weighted roots
        lo = offset
        for c, w in zip(roots, weights):
            sl = slice(lo, lo + c.size)
            lo += c.size
            if w != 0:
                pc = self._p(c)
                p = float(c @ pc)
                self.active.append((sl, float(w), c, p, pc - p * c))
        self._pdiag = None

    def _p(self, x):
        return self.plan.penalty(x, self.penalty)

    def gradient(self, g):
        out = numpy.array(g, copy=True)
        for sl, w, c, p, r in self.active:
            out[sl] += 2.0 * w * r
        return out

    def diagonal(self, hd):
        out = numpy.array(hd, copy=True)
        if self.active and self._pdiag is None:
            self._pdiag = self.penalty.diagonal(self.plan.diagonal_vector())
        for sl, w, c, p, r in self.active:
            out[sl] += 2.0 * w * (self._pdiag - p - 2.0 * r * c)
        return out

    def hessian_vector(self, base_hop):
        def hop(x):
            out = base_hop(x)
            for sl, w, c, p, r in self.active:
                v = x[sl]
                out[sl] += 2.0 * w * ((self._p(v) - p * v) - (c @ v) * r - (r @ v) * c)
            return out
        return hop

    def update(self, base_update, nroots):
        weights = {sl.start: w for sl, w, *_ in self.active}

        def upd(u, ci):
            out = base_update(u, ci)
            vecs = [ci] if nroots == 1 else list(ci)
            lo = self.offset
            for vec in vecs:
                c = numpy.asarray(vec, dtype=numpy.float64).ravel()
                w = weights.get(lo, 0.0)
                if w != 0.0:
                    c = c / numpy.linalg.norm(c)
                    pc = self._p(c)
                    out[lo:lo + c.size] += 2.0 * w * (pc - (c @ pc) * c)
                lo += c.size
            return out
        return upd


def gen_g_hop(mc, mo, ci0, eris, verbose=None, base=None):
    """(gradient, update, hop, hdiag) of the Newton step with the spin
    penalty's CI terms added (the reference's gasscf.py:213-289).  The
    orbital rows and the Hamiltonian's CI rows are the base driver's;
    ``base`` has the signature of ``newton_casscf.gen_g_hop``."""
    _check_symmetry(mc, mo)
    gradient, update, hop, hdiag = (base or _newton_base())(mc, mo, ci0, eris, verbose)
    solver = mc.fcisolver
    penalty = fci_gas.SpinPenalty.of(solver, mc.ncas, _active_electrons(mc))
    if penalty is None or not penalty.active():
        return gradient, update, hop, hdiag
    cached = getattr(solver, "cache_plans", False) and hasattr(solver, "_get_spin_plan")
    plan = solver._get_spin_plan(mc.ncas, mc.nelecas) if cached else \
        solver.make_spin_plan(mc.ncas, mc.nelecas)
    nroots = int(getattr(solver, "nroots", 1))
    roots = [numpy.asarray(c, dtype=numpy.float64).ravel()
             for c in ([ci0] if nroots == 1 else ci0)]
    weights = tuple(getattr(mc, "weights", (1.0,)))
    offset = gradient.size - sum(c.size for c in roots)
    terms = _PenaltyNewtonTerms(plan, penalty, roots, weights, offset)
    return (terms.gradient(gradient), terms.update(update, nroots),
            terms.hessian_vector(hop), terms.diagonal(hdiag))


class _NewtonPlanCache:
    """Plans one Newton run reuses, all tied to one GAS space (the
    topology): at most ``limit`` contraction plans keyed by the absorbed
    Hamiltonian's shape and SHA-256 digest, recency-ordered (a hit moves the
    plan to the back, a miss at the limit closes the front one), plus one
    RDM plan and one S^2 plan made on demand.  A new topology closes them
    all."""

    def __init__(self, limit):
        self.limit = limit
        self.topology = None
        self.space = None
        self.contract = {}          # key -> plan; dict order = recency
        self.rdm = None
        self.spin = None

    def bind(self, topology):
        if topology != self.topology:
            self.release()
            self.topology = topology

    def contract_plan(self, key, make_space, make_plan):
        plan = self.contract.pop(key, None)
        if plan is None:
            if self.space is None:
                self.space = make_space()
            while len(self.contract) >= self.limit:
                oldest = next(iter(self.contract))
                self.contract.pop(oldest).close()
            plan = make_plan(self.space)
        self.contract[key] = plan
        return plan

    def release(self):
        for plan in self.contract.values():
            plan.close()
        self.contract.clear()
        for name in ("space", "rdm"):
            obj = getattr(self, name)
            if obj is not None:
                obj.close()
                setattr(self, name, None)
        self.spin = None
        self.topology = None


def _hamiltonian_key(eri):
    arr = numpy.ascontiguousarray(eri, dtype=numpy.float64)
    return arr, (arr.shape, hashlib.sha256(arr.tobytes()).hexdigest())


def _single_state(ci):
    """Newton passes one state as ``[ci]``; more than one is not a state."""
    if isinstance(ci, (list, tuple)):
        if len(ci) != 1:
            _unsupported("multiroot CI density dispatch")
        return ci[0]
    return ci


class _GASFCISolver(fci_gas.FCISolver):
    """GAS solver of a Newton GASSCF run (the reference's gasscf.py:292-538):
    plans reused across the run's many contractions (``cache_plans``),
    Newton's ``[ci]`` lists accepted by the density and S^2 methods, and the
    optional FCI hooks native Newton probes hidden (no CAS link strings; an
    orbital rotation does not preserve a restricted space).  Copies start
    with an empty cache; an explicit ``plan=`` is checked and used for that
    call only."""

    _MAX_CONTRACT_PLANS = 3
    gen_linkstr = None
    transform_ci_for_orbital_rotation = None

    def __init__(self, mol=None, gas_orbs=None, gas_restr=None,
                 gas_restr_type=addons_gas.GAS_RESTR_SPIN_SUPERGROUP, lib=None, *,
                 cache_plans=True):
        super().__init__(mol, gas_orbs=gas_orbs, gas_restr=gas_restr,
                         gas_restr_type=gas_restr_type, lib=lib)
        self.cache_plans = bool(cache_plans)
        self._cache = _NewtonPlanCache(self._MAX_CONTRACT_PLANS)

    # views of the cache under the names callers inspect
    @property
    def _contract_plans(self):
        return self._cache.contract

    @property
    def _spin_plan(self):
        return self._cache.spin

    @property
    def _rdm_plan_cache(self):
        return self._cache.rdm

    def _topology(self, norb, nelec):
        orbs, ne, blocks = self._space_spec(norb, nelec)
        topology = (tuple(map(int, orbs)), tuple(map(int, ne)),
                    numpy.asarray(blocks, dtype=numpy.int64).tobytes(), id(self.lib))
        self._cache.bind(topology)

    def _get_contract_plan(self, eri, norb, nelec):
        arr, key = _hamiltonian_key(eri)
        self._topology(norb, nelec)
        return self._cache.contract_plan(
            key, lambda: self.make_space(norb, nelec, compress_links=True),
            lambda space: fci_gas.GasContractPlan(space, arr.copy()))

    def _get_spin_plan(self, norb, nelec):
        self._topology(norb, nelec)
        if self._cache.spin is None:
            self._cache.spin = self.make_spin_plan(norb, nelec)
        return self._cache.spin

    def _rdm_plan(self, norb, nelec, plan):
        if plan is None and self.cache_plans:
            self._topology(norb, nelec)
            if self._cache.rdm is None:
                self._cache.rdm = self.make_rdm_plan(norb, nelec)
            plan = self._cache.rdm
        return super()._rdm_plan(norb, nelec, plan)

    def close(self):
        """Close every cached plan; safe to call repeatedly."""
        self._cache.release()

    def contract_2e(self, eri, fcivec, norb, nelec, link_index=None, *args, **kwargs):
        """P H P c; without an explicit plan, a cached plan per Hamiltonian."""
        compress = kwargs.pop("compress_links", True)
        if args or kwargs or not (self.cache_plans and compress):
            return super().contract_2e(eri, fcivec, norb, nelec, link_index, *args, **kwargs)
        plan = self._get_contract_plan(eri, norb, nelec)
        return super().contract_2e(eri, fcivec, norb, nelec, link_index, plan=plan)

    def contract_ss(self, fcivec, norb, nelec):
        fcivec = _single_state(fcivec)
        if self.cache_plans:
            return self._get_spin_plan(norb, nelec).contract(fcivec)
        return super().contract_ss(fcivec, norb, nelec)

    def spin_square(self, ci, norb, nelec, *args, **kwargs):
        return super().spin_square(_single_state(ci), norb, nelec)

    def copy(self):
        twin = _copy.copy(self)
        twin._cache = _NewtonPlanCache(self._MAX_CONTRACT_PLANS)
        return twin


def _accept_state_lists(name, nvec):
    """A density method of the base solver whose first ``nvec`` arguments
    may be Newton's one-element lists."""
    inner = getattr(fci_gas.FCISolver, name)

    def method(self, *args, **kwargs):
        args = tuple(_single_state(a) for a in args[:nvec]) + args[nvec:]
        return inner(self, *args, **kwargs)
    method.__name__ = name
    method.__doc__ = inner.__doc__
    return method


for _name, _nvec in (("make_rdm1s", 1), ("make_rdm1", 1), ("make_rdm12s", 1),
                     ("make_rdm12", 1), ("trans_rdm1s", 2), ("trans_rdm1", 2),
                     ("trans_rdm12s", 2), ("trans_rdm12", 2)):
    setattr(_GASFCISolver, _name, _accept_state_lists(_name, _nvec))
del _name, _nvec


def _gas_orbs(spec):
    values = numpy.asarray(spec).reshape(-1)
    orbs = tuple(int(v) for v in values)
    if not orbs or min(orbs) <= 0 or not numpy.array_equal(values.astype(float), orbs):
        raise ValueError("gas_orbs must be positive integers")
    return orbs


def new_gas_solver(mf, gas_orbs, gas_restr=None, gas_restr_type=None, cache_plans=None):
    """The solver of a GASSCF object built from a mean field
    (gasscf.py:547-558)."""
    if gas_orbs is None:
        raise ValueError("gas_orbs is required without an explicit GASCI solver")
    return _GASFCISolver(getattr(mf, "mol", None), gas_orbs=_gas_orbs(gas_orbs),
                         gas_restr=gas_restr,
                         gas_restr_type=gas_restr_type or addons_gas.GAS_RESTR_SPIN_SUPERGROUP,
                         cache_plans=True if cache_plans is None else bool(cache_plans))


def adapt_solver(fcisolver, cache_plans=None):
    """A Newton solver made from an explicit GAS solver (gasscf.py:572-600):
    its settings, a cache of its own."""
    if isinstance(fcisolver, _GASFCISolver):
        solver = fcisolver.copy()
    elif type(fcisolver) is fci_gas.FCISolver:
        solver = _copy.copy(fcisolver)
        solver.__class__ = _GASFCISolver
        solver._cache = _NewtonPlanCache(_GASFCISolver._MAX_CONTRACT_PLANS)
    else:
        _unsupported("external/non-GASCI solver adaptation")
    if solver.gas_orbs is None:
        raise ValueError("explicit GASCI solver must define gas_orbs")
    solver.gas_orbs = _gas_orbs(solver.gas_orbs)
    solver.gas_restr_type = solver.gas_restr_type or addons_gas.GAS_RESTR_SPIN_SUPERGROUP
    solver.cache_plans = bool(getattr(fcisolver, "cache_plans", True)
                              if cache_plans is None else cache_plans)
    return solver

"""Drop-in direct-CI solver on MI355X with pyscf's ``direct_spin1`` API.

Module functions and ``FCISolver`` keep the signatures of
pyscf.fci.direct_spin1 as the reference's solvers call them (the call
sites: ``absorb_h1e(h1e, eri, norb, nelec, fac=1)`` csf.py:518,
``contract_2e(eri, fcivec, norb, nelec, link_index=None)`` fci_gas.py:919 /
csf.py:525, ``make_hdiag`` fci_gas.py:904, ``kernel`` fci_gas.py:1150,
``make_rdm1(s)/make_rdm12(s)/trans_rdm12(s)`` fci_gas.py:1401-1483,
``energy`` fci_gas.py:946).  CAS CI vectors are (nstr_alpha, nstr_beta)
float64 arrays.  Host numpy inputs are returned as numpy; device torch
tensors stay on the device.

Every call runs the HIP library (libfcisigma.so); there is no CPU path.
"""

import ctypes
import math

import numpy

from . import _lib
from . import davidson as _davidson

PSPACE_MATVEC_MAX = 64   # exact small-space diagonalisation bound (fci_gas.py:38)
INIT_NOISE_NORM = 1e-3   # fci_gas.py:40


def _unpack_nelec(nelec, spin=None):
    if isinstance(nelec, (int, numpy.integer)):
        if spin is None:
            nb = int(nelec) // 2
        else:
            nb = (int(nelec) - int(spin)) // 2
        na = int(nelec) - nb
        return na, nb
    return int(nelec[0]), int(nelec[1])


def pair_index(p, q):
    hi, lo = max(p, q), min(p, q)
    return hi * (hi + 1) // 2 + lo


def pair_matrix(eri, norb):
    """(npair, npair) float64 4-fold pair matrix from a pair matrix, a
    4-index tensor or an 8-fold packed vector (the layouts
    _as_pair_matrix accepts, pyscf/mcscf/fci_gas.py:101-124)."""
    norb = int(norb)
    npair = norb * (norb + 1) // 2
    arr = numpy.asarray(eri)
    if numpy.iscomplexobj(arr):
        raise TypeError("the sigma kernels require real-valued integrals")
    arr = numpy.asarray(arr, dtype=numpy.float64)
    if arr.shape == (npair, npair):
        return numpy.ascontiguousarray(arr)
    if arr.shape == (norb,) * 4:
        i, j = numpy.tril_indices(norb)
        return numpy.ascontiguousarray(arr[i, j][:, i, j])
    if arr.ndim == 1 and arr.size == npair * (npair + 1) // 2:
        out = numpy.empty((npair, npair))
        iu = numpy.tril_indices(npair)
        out[iu] = arr
        out.T[iu] = arr
        return out
    raise ValueError("could not convert ERI of shape %s to pair layout" % (arr.shape,))


def absorb_h1e(h1e, eri, norb, nelec, fac=1):
    """Absorbed two-electron Hamiltonian (pair layout); the reference's
    solvers call it with fac=0.5 (fci_gas.py:1166, csf.py:445)."""
    lib = _lib.load()
    norb = int(norb)
    ne = sum(_unpack_nelec(nelec))
    h1e = numpy.ascontiguousarray(h1e, dtype=numpy.float64).reshape(norb, norb)
    eri = pair_matrix(eri, norb)
    out = numpy.empty_like(eri)
    _lib.check(lib.fcis_absorb_h1e(norb, ne, float(fac), _lib.dptr(h1e),
                                   _lib.dptr(eri), _lib.dptr(out)), "fcis_absorb_h1e")
    return out


# ---------------------------------------------------------------- spaces

class Space:
    """Device string space (link tables built on the GPU).  Addresses are
    colex ranks (pyscf's order) unless ``strings`` = (alpha, beta) uint64
    arrays give another address order (fcis_space_create_ordered)."""

    def __init__(self, norb, nelec, strings=None):
        import ctypes
        self.lib = _lib.load()
        self.norb = int(norb)
        self.nelec = _unpack_nelec(nelec)
        self.ordered = strings is not None
        h = ctypes.c_void_p()
        if strings is None:
            _lib.check(self.lib.fcis_space_create(ctypes.byref(h), self.norb, self.nelec[0],
                                                  self.nelec[1], _lib.stream_ptr()),
                       "fcis_space_create")
        else:
            sa, sb = (numpy.ascontiguousarray(x, dtype=numpy.uint64) for x in strings)
            up = ctypes.POINTER(ctypes.c_uint64)
            _lib.check(self.lib.fcis_space_create_ordered(
                ctypes.byref(h), self.norb, self.nelec[0], self.nelec[1],
                sa.ctypes.data_as(up), sb.ctypes.data_as(up), _lib.stream_ptr()),
                "fcis_space_create_ordered")
        self.handle = h
        na, nb = ctypes.c_int64(), ctypes.c_int64()
        la, lb = ctypes.c_int(), ctypes.c_int()
        _lib.check(self.lib.fcis_space_info(h, ctypes.byref(na), ctypes.byref(nb),
                                            ctypes.byref(la), ctypes.byref(lb)),
                   "fcis_space_info")
        self.nstra, self.nstrb = na.value, nb.value
        self.nlinka, self.nlinkb = la.value, lb.value
        self.ndet = self.nstra * self.nstrb
        self.shape = (self.nstra, self.nstrb)

    def link_index(self, spin=0, tril=False):
        """pyscf gen_linkstr_index layout [nstr, nlink, 4] (device tensor)."""
        import torch
        nstr = self.nstra if spin == 0 else self.nstrb
        nlink = self.nlinka if spin == 0 else self.nlinkb
        out = torch.zeros((nstr, nlink, 4), dtype=torch.int32, device="cuda")
        _lib.check(self.lib.fcis_gen_linkstr_index(self.handle, int(spin), 1 if tril else 0,
                                                   _lib.vptr(out), _lib.stream_ptr()),
                   "fcis_gen_linkstr_index")
        return out

    def close(self):
        if getattr(self, "handle", None) is not None:
            self.lib.fcis_space_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_SPACES = {}


def get_space(norb, nelec):
    key = (int(norb),) + _unpack_nelec(nelec)
    sp = _SPACES.get(key)
    if sp is None:
        sp = _SPACES[key] = Space(norb, key[1:])
    return sp


def gen_linkstr_index(orb_list, nocc, tril=False):
    """cistring.gen_linkstr_index for range(norb), built on the device,
    returned as a host int32 array [nstr, nlink, 4]."""
    norb = len(orb_list) if hasattr(orb_list, "__len__") else int(orb_list)
    sp = get_space(norb, (nocc, nocc))
    return sp.link_index(0, tril).cpu().numpy()


class SigmaPlan:
    """Reusable sigma plan (fcis_sigma_plan_*), the analogue of the
    reference's GasContractPlan (pyscf/mcscf/fci_gas.py:570-650)."""

    def __init__(self, space, h2e, precision="fp64", workspace_bytes=0):
        import ctypes
        self.lib = _lib.load()
        self.space = space
        prec = _lib.PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        self.precision = precision
        self.h2e = pair_matrix(h2e, space.norb)
        h = ctypes.c_void_p()
        _lib.check(self.lib.fcis_sigma_plan_create(ctypes.byref(h), space.handle,
                                                   _lib.dptr(self.h2e), prec,
                                                   int(workspace_bytes)),
                   "fcis_sigma_plan_create")
        self.handle = h
        nt, ta, ws = ctypes.c_int(), ctypes.c_int64(), ctypes.c_uint64()
        self.lib.fcis_sigma_plan_info(h, ctypes.byref(nt), ctypes.byref(ta), ctypes.byref(ws))
        self.ntiles, self.tile_alpha, self.workspace_bytes = nt.value, ta.value, ws.value
        self.fused = self.lib.fcis_sigma_plan_path(h) == 1

    def __call__(self, c, out):
        """out <- H c (device float64 tensors with space.ndet elements)."""
        if c.data_ptr() == out.data_ptr():
            raise ValueError("sigma input and output must not alias")
        _lib.check(self.lib.fcis_sigma_plan_execute(self.handle, _lib.vptr(c), _lib.vptr(out),
                                                    _lib.stream_ptr()),
                   "fcis_sigma_plan_execute")
        return out

    def set_timing(self, enable=True):
        """Per-kernel HIP-event timing of gather / contraction / scatter."""
        _lib.check(self.lib.fcis_sigma_plan_set_timing(self.handle, 1 if enable else 0),
                   "fcis_sigma_plan_set_timing")

    def timing(self):
        """(ms[3], launches[3]) accumulated since set_timing(True)."""
        ms = numpy.zeros(3)
        n = (ctypes.c_int64 * 3)()
        _lib.check(self.lib.fcis_sigma_plan_get_timing(self.handle, _lib.dptr(ms), n),
                   "fcis_sigma_plan_get_timing")
        return ms, numpy.array(list(n))

    def close(self):
        if getattr(self, "handle", None) is not None:
            self.lib.fcis_sigma_plan_free(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _to_device(x, ndet):
    import torch
    if isinstance(x, torch.Tensor):
        t = x.detach().to(device="cuda", dtype=torch.float64).contiguous().reshape(-1)
        host = False
    else:
        a = numpy.ascontiguousarray(x, dtype=numpy.float64).reshape(-1)
        t = torch.from_numpy(a).to("cuda")
        host = True
    if t.numel() != ndet:
        raise ValueError("CI vector size %d does not match determinant count %d"
                         % (t.numel(), ndet))
    return t, host


def available_bytes():
    """Free device memory plus the plan workspaces the library keeps for
    reuse (fcis_kept_memory), which a new plan takes over."""
    import ctypes
    import torch
    free, _ = torch.cuda.mem_get_info()
    kept = ctypes.c_uint64()
    _lib.check(_lib.load().fcis_kept_memory(torch.cuda.current_device(), ctypes.byref(kept)),
               "fcis_kept_memory")
    return int(free) + int(kept.value)


def _default_workspace():
    return int(available_bytes() * 0.45)


# Plans of the module-level contract_2e, reused across calls with the same
# space, precision and absorbed Hamiltonian (the reference's solvers call
# contract_2e once per Davidson vector with unchanged integrals, and its
# GasContractPlan is built once per Hamiltonian, fci_gas.py:570-650).  A plan
# holds device tables and the sigma workspace, so at most PLAN_CACHE_SIZE are
# kept, and a plan whose workspace exceeds PLAN_CACHE_FRACTION of the device
# memory is freed after its call; release_plans() frees them all.
PLAN_CACHE_SIZE = 2
PLAN_CACHE_FRACTION = 0.25
_PLANS = {}            # key -> SigmaPlan, dict order = recency


def release_plans():
    """Free the plans cached by contract_2e and the plan workspaces the
    library keeps for reuse."""
    while _PLANS:
        _PLANS.pop(next(iter(_PLANS))).close()
    _lib.check(_lib.load().fcis_release_kept_memory(-1), "fcis_release_kept_memory")


def _cached_plan(sp, eri, precision, workspace_bytes):
    import hashlib
    h = pair_matrix(eri, sp.norb)
    key = (sp.norb, sp.nelec, str(precision), int(workspace_bytes),
           hashlib.sha256(h.tobytes()).hexdigest())
    plan = _PLANS.pop(key, None)
    if plan is None:
        while len(_PLANS) >= PLAN_CACHE_SIZE:
            _PLANS.pop(next(iter(_PLANS))).close()
        try:
            plan = SigmaPlan(sp, h, precision, workspace_bytes or _default_workspace())
        except RuntimeError:
            # the cached plans' workspaces may be what is missing
            release_plans()
            plan = SigmaPlan(sp, h, precision, workspace_bytes or _default_workspace())
    _PLANS[key] = plan
    return plan


def contract_2e(eri, fcivec, norb, nelec, link_index=None, precision="fp64", **kwargs):
    """sigma = H c for the absorbed pair matrix ``eri`` (direct_spin1.contract_2e).

    Host (numpy) vectors are copied to the device and back; device tensors
    stay there.  The plan for (space, precision, eri) is kept for the next
    call (release_plans frees it)."""
    sp = get_space(norb, nelec)
    c, host = _to_device(fcivec, sp.ndet)
    import torch
    out = torch.empty_like(c)
    plan = _cached_plan(sp, eri, precision, int(kwargs.get("workspace_bytes", 0) or 0))
    plan(c, out)
    if plan.workspace_bytes > PLAN_CACHE_FRACTION * torch.cuda.mem_get_info()[1]:
        for key in [k for k, v in _PLANS.items() if v is plan]:
            _PLANS.pop(key).close()
    shape = numpy.shape(fcivec) if host else tuple(fcivec.shape)
    if host:
        return out.cpu().numpy().reshape(shape)
    return out.reshape(shape)


def make_hdiag(h1e, eri, norb, nelec, compress=False):
    """Hamiltonian diagonal (un-absorbed integrals), device tensor if
    ``device=True`` else numpy (direct_spin1.make_hdiag)."""
    return _make_hdiag_dev(h1e, eri, norb, nelec).cpu().numpy()


def _make_hdiag_dev(h1e, eri, norb, nelec):
    import torch
    sp = get_space(norb, nelec)
    h1e = numpy.ascontiguousarray(h1e, dtype=numpy.float64).reshape(sp.norb, sp.norb)
    eri = pair_matrix(eri, sp.norb)
    out = torch.empty(sp.ndet, dtype=torch.float64, device="cuda")
    _lib.check(sp.lib.fcis_make_hdiag(sp.handle, _lib.dptr(h1e), _lib.dptr(eri),
                                      _lib.vptr(out), _lib.stream_ptr()), "fcis_make_hdiag")
    return out


def _pspace_dev(sp, h1e, eri, addr):
    """H block over the CAS determinants ``addr`` (device int64 tensor) by
    the Slater-Condon rules (fcis_pspace), as a device (m, m) tensor.  A
    two-component h1e (h_alpha, h_beta) gives the UHF block
    (fcis_pspace_uhf)."""
    import torch
    h = numpy.asarray(h1e, dtype=numpy.float64)
    eri = pair_matrix(eri, sp.norb)
    addr = addr.to(device="cuda", dtype=torch.int64).contiguous()
    m = addr.numel()
    h0 = torch.empty((m, m), dtype=torch.float64, device="cuda")
    if h.ndim == 3:
        ha = numpy.ascontiguousarray(h[0]).reshape(sp.norb, sp.norb)
        hb = numpy.ascontiguousarray(h[1]).reshape(sp.norb, sp.norb)
        _lib.check(sp.lib.fcis_pspace_uhf(sp.handle, _lib.dptr(ha), _lib.dptr(hb), _lib.dptr(eri),
                                          int(m), _lib.vptr(addr), _lib.vptr(h0),
                                          _lib.stream_ptr()), "fcis_pspace_uhf")
        return h0
    h = numpy.ascontiguousarray(h).reshape(sp.norb, sp.norb)
    _lib.check(sp.lib.fcis_pspace(sp.handle, _lib.dptr(h), _lib.dptr(eri), int(m),
                                  _lib.vptr(addr), _lib.vptr(h0), _lib.stream_ptr()),
               "fcis_pspace")
    return h0


def contract_1e_ab_dev(sp, h1a, h1b, x, out, accumulate=False):
    """out (+)= sum_pq h1a[p,q] E^a_pq x + h1b[p,q] E^b_pq x on device
    vectors (fcis_contract_1e_ab); h1a / h1b host or device (norb, norb)."""
    import torch

    def dev(h):
        return torch.as_tensor(numpy.asarray(h, dtype=numpy.float64) if not isinstance(
            h, torch.Tensor) else h).to(device="cuda", dtype=torch.float64).contiguous()
    ha, hb = dev(h1a), dev(h1b)
    if ha.numel() != sp.norb ** 2 or hb.numel() != sp.norb ** 2:
        raise ValueError("one-electron arrays must be norb x norb")
    _lib.check(sp.lib.fcis_contract_1e_ab(sp.handle, _lib.vptr(ha), _lib.vptr(hb), _lib.vptr(x),
                                          _lib.vptr(out), 1 if accumulate else 0,
                                          _lib.stream_ptr()), "fcis_contract_1e_ab")
    return out


def pspace(h1e, eri, norb, nelec, hdiag=None, np=400):
    """(addr, h0): the ``np`` determinants of lowest diagonal energy
    (ascending) and H over them, from the un-absorbed integrals
    (direct_spin1.pspace; built on the GPU by the Slater-Condon rules)."""
    import torch
    sp = get_space(norb, nelec)
    hd = _make_hdiag_dev(h1e, eri, norb, nelec) if hdiag is None else \
        torch.as_tensor(numpy.asarray(hdiag, dtype=numpy.float64).reshape(-1)).cuda()
    if hd.numel() != sp.ndet:
        raise ValueError("hdiag size does not match the determinant count")
    m = min(int(np), sp.ndet)
    addr = lowest_k(hd, m)
    h0 = _pspace_dev(sp, h1e, eri, addr)
    return addr.cpu().numpy(), h0.cpu().numpy()


class SpinPlan:
    """S^2 on device CAS vectors (fcis_spin_plan_*): contract(c, out) and
    diag(out).  The space must outlive the plan."""

    def __init__(self, space):
        import ctypes
        self.lib = _lib.load()
        self.space = space
        h = ctypes.c_void_p()
        _lib.check(self.lib.fcis_spin_plan_create(ctypes.byref(h), space.handle,
                                                  _lib.stream_ptr()), "fcis_spin_plan_create")
        self.handle = h

    def contract(self, c, out):
        if c.data_ptr() == out.data_ptr():
            raise ValueError("S^2 input and output must not alias")
        _lib.check(self.lib.fcis_spin_plan_contract(self.handle, _lib.vptr(c), _lib.vptr(out),
                                                    _lib.stream_ptr()), "fcis_spin_plan_contract")
        return out

    def diag(self, out):
        _lib.check(self.lib.fcis_spin_plan_diag(self.handle, _lib.vptr(out), _lib.stream_ptr()),
                   "fcis_spin_plan_diag")
        return out

    def close(self):
        if getattr(self, "handle", None) is not None:
            self.lib.fcis_spin_plan_free(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def contract_ss(fcivec, norb, nelec):
    """S^2 c (pyscf.fci.spin_op.contract_ss; the reference's S^2 contraction,
    pyscf/mcscf/fci_gas.py:346-372), numpy in / numpy out or device tensors."""
    import torch
    sp = get_space(norb, nelec)
    c, host = _to_device(fcivec, sp.ndet)
    out = torch.empty_like(c)
    with SpinPlan(sp) as plan:
        plan.contract(c, out)
    shape = numpy.shape(fcivec) if host else tuple(fcivec.shape)
    return out.cpu().numpy().reshape(shape) if host else out.reshape(shape)


def trans_rdm12s(cibra, ciket, norb, nelec, link_index=None, reorder=True):
    """((dm1a, dm1b), (dm2aa, dm2ab, dm2ba, dm2bb)) as direct_spin1.trans_rdm12s."""
    (dm1a, dm1b), (dm2aa, dm2ab, dm2bb) = _rdm12s(cibra, ciket, norb, nelec)
    _, (_, dm2ba_ji, _) = _rdm12s(ciket, cibra, norb, nelec)
    return (dm1a, dm1b), (dm2aa, dm2ab, dm2ba_ji.transpose(3, 2, 1, 0), dm2bb)


def _rdm12s(cibra, ciket, norb, nelec, with_dm2=True):
    sp = get_space(norb, nelec)
    bra, _ = _to_device(cibra, sp.ndet)
    ket, _ = _to_device(ciket, sp.ndet)
    n = sp.norb
    d1a, d1b = numpy.empty((n, n)), numpy.empty((n, n))
    if with_dm2:
        d2 = [numpy.empty((n,) * 4) for _ in range(3)]
        ptrs = [_lib.dptr(x) for x in d2]
    else:
        d2 = [None] * 3
        ptrs = [None] * 3
    _lib.check(sp.lib.fcis_trans_rdm12s(sp.handle, _lib.vptr(bra), _lib.vptr(ket),
                                        _lib.dptr(d1a), _lib.dptr(d1b), *ptrs,
                                        _lib.stream_ptr()), "fcis_trans_rdm12s")
    return (d1a, d1b), tuple(d2)


def make_rdm12s(fcivec, norb, nelec, link_index=None, reorder=True):
    """((dm1a, dm1b), (dm2aa, dm2ab, dm2bb)) (direct_spin1.make_rdm12s)."""
    return _rdm12s(fcivec, fcivec, norb, nelec)


def make_rdm12(fcivec, norb, nelec, link_index=None, reorder=True):
    """Spin-traced (dm1, dm2), dm2 = aa + bb + ab + ab^T(2,3,0,1)
    (assembly of pyscf/mcscf/fci_gas.py:724-733)."""
    (dm1a, dm1b), (dm2aa, dm2ab, dm2bb) = make_rdm12s(fcivec, norb, nelec)
    dm2 = dm2aa + dm2bb + dm2ab + dm2ab.transpose(2, 3, 0, 1)
    return dm1a + dm1b, dm2


def trans_rdm12(cibra, ciket, norb, nelec, link_index=None, reorder=True):
    (dm1a, dm1b), (dm2aa, dm2ab, dm2bb) = _rdm12s(cibra, ciket, norb, nelec)
    _, (_, dm2ba_ji, _) = _rdm12s(ciket, cibra, norb, nelec)
    dm2ba = dm2ba_ji.transpose(3, 2, 1, 0)
    return dm1a + dm1b, dm2aa + dm2ab + dm2ba + dm2bb


def make_rdm1s(fcivec, norb, nelec, link_index=None):
    (dm1a, dm1b), _ = _rdm12s(fcivec, fcivec, norb, nelec, with_dm2=False)
    return dm1a, dm1b


def make_rdm1(fcivec, norb, nelec, link_index=None):
    dm1a, dm1b = make_rdm1s(fcivec, norb, nelec)
    return dm1a + dm1b


def trans_rdm1s(cibra, ciket, norb, nelec, link_index=None):
    (dm1a, dm1b), _ = _rdm12s(cibra, ciket, norb, nelec, with_dm2=False)
    return dm1a, dm1b


def trans_rdm1(cibra, ciket, norb, nelec, link_index=None):
    dm1a, dm1b = trans_rdm1s(cibra, ciket, norb, nelec)
    return dm1a + dm1b


def energy(h1e, eri, fcivec, norb, nelec, link_index=None):
    """<c|H|c> for a normalised c (direct_spin1.energy / fci_gas.py:946-952)."""
    import torch
    sp = get_space(norb, nelec)
    h2e = absorb_h1e(h1e, eri, norb, nelec, 0.5)
    c, _ = _to_device(fcivec, sp.ndet)
    out = torch.empty_like(c)
    with SigmaPlan(sp, h2e, "fp64", _default_workspace()) as plan:
        plan(c, out)
    return float(torch.dot(c, out))


def spin_square_from_rdm12s(dm1s, dm2s, nelec):
    """<S^2> from spin-resolved RDMs (pyscf/mcscf/fci_gas.py:1486-1499)."""
    na, nb = _unpack_nelec(nelec)
    dm2ab = dm2s[1]
    sz = 0.5 * (na - nb)
    ss = sz * (sz + 1.0) + nb - float(numpy.einsum("pqqp->", dm2ab))
    return ss, max(0.0, 4.0 * ss + 1.0) ** 0.5


# ---------------------------------------------------------------- solver

LOWEST_K_CHUNK = 16384


def lowest_k(hd, k):
    """Indices of the k lowest entries of a 1-D tensor, by ascending value
    (torch.topk(hd, k, largest=False).indices).  On long vectors: the k lowest
    entries lie in the k chunks of LOWEST_K_CHUNK entries with the lowest
    minima (the k-th lowest chunk minimum bounds the k-th lowest entry), so
    one min-reduction over the vector and a topk over k chunks (plus the
    tail) replace a selection over all of it."""
    import torch
    n, k = hd.numel(), int(k)
    C = LOWEST_K_CHUNK
    m = n // C
    if k < 1 or m < 4 * k:
        return torch.topk(hd, k, largest=False).indices
    body = hd[:m * C].view(m, C)
    rows = torch.topk(body.amin(dim=1), k, largest=False).indices
    cols = torch.arange(C, device=hd.device, dtype=rows.dtype)
    cidx = torch.cat([(rows[:, None] * C + cols[None, :]).reshape(-1),
                      torch.arange(m * C, n, device=hd.device, dtype=rows.dtype)])
    sel = torch.topk(hd[cidx], k, largest=False).indices
    return cidx[sel]


def lowest_addresses(hd, count):
    """Addresses of the ``count`` lowest entries of a device diagonal, in the
    order of the reference's _lowest_hdiag_addresses (pyscf/mcscf/fci_gas.py:
    46-58): the whole range 0..n-1 when count >= n, else
    numpy.argpartition(hd, count - 1)[:count], which lists them by ascending
    value whenever the count + 1 lowest values are distinct.  Ties (the
    alpha <-> beta mirror determinants of spaces with na == nb) are broken
    here by the lower address; numpy's own tie order depends on the
    selection routine its build dispatches to.  Host int64 array."""
    import torch
    n = hd.numel()
    count = int(count)
    if count >= n:
        return numpy.arange(n, dtype=numpy.int64)
    if count == 1:
        # one reduction: argmin returns the first (lowest-address) minimum
        return numpy.array([int(torch.argmin(hd))], dtype=numpy.int64)
    # the count-th lowest value, then every entry at or below it (ties at
    # the boundary included) ordered by (value, address) on the host
    thr = torch.topk(hd, count, largest=False, sorted=True).values[-1]
    cand = torch.nonzero(hd <= thr).reshape(-1)
    val = hd[cand].cpu().numpy()
    cand = cand.cpu().numpy()
    return cand[numpy.lexsort((cand, val))[:count]].astype(numpy.int64)


class FCISolver:
    """CAS FCI solver with pyscf's fcisolver protocol (kernel, contract_2e,
    absorb_h1e, make_hdiag, make_rdm1(s)/make_rdm12(s)/trans_rdm*, energy,
    spin_square), running every O(N_det) step on the GPU.

    Extra attribute ``precision`` selects the sigma arithmetic used during
    the Davidson iterations ("fp64", "fp32", "bf16x3"); non-FP64 solves end
    with FP64 sigma refinement (see ``mixed``)."""

    def __init__(self, mol=None):
        self.mol = mol
        self.verbose = 0
        self.max_memory = 4000
        self.nroots = 1
        self.spin = None
        self.conv_tol = 1e-10
        self.conv_tol_residual = None
        self.max_cycle = 50
        self.max_space = 12
        self.lindep = 1e-14
        self.level_shift = 1e-3
        self.davidson_only = False
        self.pspace_size = 400
        self.precision = "fp64"
        self.workspace_bytes = 0
        self.converged = False
        self.eci = None
        self.ci = None
        self.norb = None
        self.nelec = None
        self.stats = {}

    # -- pyscf protocol -------------------------------------------------
    def absorb_h1e(self, h1e, eri, norb, nelec, fac=1):
        return absorb_h1e(h1e, eri, norb, nelec, fac)

    def contract_2e(self, eri, fcivec, norb, nelec, link_index=None, **kwargs):
        return contract_2e(eri, fcivec, norb, nelec, link_index,
                           precision=kwargs.pop("precision", "fp64"))

    def make_hdiag(self, h1e, eri, norb, nelec, *args, **kwargs):
        return make_hdiag(h1e, eri, norb, nelec)

    def energy(self, h1e, eri, fcivec, norb, nelec, link_index=None):
        return energy(h1e, eri, fcivec, norb, nelec)

    def make_rdm1s(self, ci, norb, nelec, link_index=None):
        return make_rdm1s(ci, norb, self._nelec(nelec))

    def make_rdm1(self, ci, norb, nelec, link_index=None):
        return make_rdm1(ci, norb, self._nelec(nelec))

    def make_rdm12s(self, ci, norb, nelec, link_index=None, reorder=True):
        return make_rdm12s(ci, norb, self._nelec(nelec))

    def make_rdm12(self, ci, norb, nelec, link_index=None, reorder=True):
        return make_rdm12(ci, norb, self._nelec(nelec))

    def trans_rdm1s(self, cibra, ciket, norb, nelec, link_index=None):
        return trans_rdm1s(cibra, ciket, norb, self._nelec(nelec))

    def trans_rdm1(self, cibra, ciket, norb, nelec, link_index=None):
        return trans_rdm1(cibra, ciket, norb, self._nelec(nelec))

    def trans_rdm12s(self, cibra, ciket, norb, nelec, link_index=None, reorder=True):
        return trans_rdm12s(cibra, ciket, norb, self._nelec(nelec))

    def trans_rdm12(self, cibra, ciket, norb, nelec, link_index=None, reorder=True):
        return trans_rdm12(cibra, ciket, norb, self._nelec(nelec))

    def spin_square(self, ci, norb, nelec, *args, **kwargs):
        nelec = self._nelec(nelec)
        dm1s, dm2s = make_rdm12s(ci, norb, nelec)
        return spin_square_from_rdm12s(dm1s, dm2s, nelec)

    def contract_ss(self, fcivec, norb, nelec):
        return contract_ss(fcivec, norb, self._nelec(nelec))

    def _nelec(self, nelec):
        return _unpack_nelec(nelec, self.spin)

    def pspace(self, h1e, eri, norb, nelec, hdiag=None, np=400):
        return pspace(h1e, eri, norb, self._nelec(nelec), hdiag, np)

    def get_init_guess(self, norb, nelec, nroots, hdiag):
        """Lowest-diagonal unit vectors plus reproducible global noise
        (fci_gas.py:881-902), as host arrays."""
        return [g.cpu().numpy() for g in self._init_guess_dev(hdiag, nroots)]

    def _init_guess_dev(self, hdiag, nroots):
        import torch
        hd = hdiag if isinstance(hdiag, torch.Tensor) else torch.as_tensor(
            numpy.asarray(hdiag, dtype=numpy.float64)).cuda()
        n = hd.numel()
        addrs = lowest_addresses(hd, nroots)
        lib = _lib.load()
        out = []
        for root, a in enumerate(addrs):
            g = torch.empty(n, dtype=torch.float64, device=hd.device)
            _lib.check(lib.fcis_init_guess(n, int(a), root, INIT_NOISE_NORM, _lib.vptr(g),
                                           _lib.stream_ptr()), "fcis_init_guess")
            out.append(g)
        return out

    def kernel(self, h1e, eri, norb, nelec, ci0=None, ecore=0, nroots=None, **kwargs):
        """Davidson solve; returns (e + ecore, ci) like direct_spin1.kernel."""
        import torch
        nroots = int(nroots if nroots is not None else kwargs.pop("nroots", self.nroots))
        if nroots < 1:
            raise ValueError("nroots must be positive")

        def option(name, default):
            v = kwargs.pop(name, None)
            return default if v is None else v

        tol = option("tol", self.conv_tol)
        tol_residual = option("tol_residual", self.conv_tol_residual)
        max_cycle = option("max_cycle", self.max_cycle)
        max_space = option("max_space", self.max_space)
        lindep = option("lindep", self.lindep)
        level_shift = option("level_shift", self.level_shift)
        davidson_only = bool(option("davidson_only", self.davidson_only))
        pspace_size = int(option("pspace_size", self.pspace_size))
        precision = option("precision", self.precision)
        verbose = option("verbose", self.verbose)
        nelec = self._nelec(nelec)
        self.norb, self.nelec = int(norb), nelec
        sp = get_space(norb, nelec)
        if nroots > sp.ndet:
            raise ValueError("nroots exceeds the determinant count")
        h1e = numpy.ascontiguousarray(h1e, dtype=numpy.float64).reshape(sp.norb, sp.norb)
        eri = pair_matrix(eri, sp.norb)
        h2e = absorb_h1e(h1e, eri, norb, nelec, 0.5)
        hdiag = _make_hdiag_dev(h1e, eri, norb, nelec)

        if (ci0 is None and not davidson_only and sp.ndet <= min(PSPACE_MATVEC_MAX,
                                                                 pspace_size)):
            e, c = self._exact_small(sp, h1e, eri, nroots)
            conv = numpy.ones(nroots, dtype=bool)
        else:
            from .mixed import solve
            from .davidson import make_pspace_precond
            # pyscf's direct_spin1 kernel preconditions with the exact block
            # of the pspace_size lowest-diagonal determinants (Olsen's
            # correction); pspace_size = 0 gives the diagonal preconditioner
            precond = None
            npsp = min(int(pspace_size), sp.ndet)
            if npsp > 0 and sp.ndet > PSPACE_MATVEC_MAX:
                addr = lowest_k(hdiag, npsp)
                pw, pv = numpy.linalg.eigh(_pspace_dev(sp, h1e, eri, addr).cpu().numpy())
                precond = make_pspace_precond(hdiag, pw, pv, addr, level_shift)
            e, c, conv, stats = solve(self, sp, h2e, hdiag, nroots, ci0, precision,
                                      tol=tol, tol_residual=tol_residual,
                                      max_cycle=max_cycle, max_space=max_space,
                                      lindep=lindep, level_shift=level_shift,
                                      verbose=verbose, precond=precond)
            self.stats = stats
        e = numpy.asarray(e, dtype=numpy.float64).reshape(-1) + float(ecore)
        cis = [x.reshape(sp.shape) for x in c]
        if nroots == 1:
            self.converged = bool(conv[0])
            self.eci = float(e[0])
            self.ci = cis[0].cpu().numpy()
        else:
            self.converged = numpy.asarray(conv, dtype=bool)
            self.eci = e
            self.ci = [x.cpu().numpy() for x in cis]
        return self.eci, self.ci

    def _exact_small(self, sp, h1e, eri, nroots):
        """Dense H of the whole space (Slater-Condon block on the GPU), then
        eigh (the reference's small-space path, fci_gas.py:1226-1257)."""
        import torch
        addr = torch.arange(sp.ndet, dtype=torch.int64, device="cuda")
        Hh = _pspace_dev(sp, h1e, eri, addr).cpu().numpy()
        w, v = numpy.linalg.eigh(Hh)
        vecs = [torch.from_numpy(numpy.ascontiguousarray(v[:, k])).cuda() for k in range(nroots)]
        return w[:nroots], vecs

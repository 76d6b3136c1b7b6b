"""Restricted (GAS) CI on the device, with the interface of pyscf-forge's
GAS solver (pyscf/mcscf/fci_gas.py).

CI vectors are one-dimensional arrays in the reference's canonical GAS block
order.  The restricted Hamiltonian is P H P, P the projection on the GAS
determinants (the identity the reference's test checks libfci_gas against,
pyscf/mcscf/test/test_gasci.py:522-575): every sigma embeds the GAS vector
in the CAS array on the device (gas2fci), runs the CAS sigma kernels and
projects back (fci2gas); the diagonal and the RDMs come from the CAS kernels
the same way.  Up to 16 orbitals the fused CAS kernel runs only over the
determinants within one excitation of the space.  Above that (and on
request, ``native=True``) the space is sector-native: the sigma, diagonal,
RDMs and H blocks run on the GAS determinants (the sigma through one-spin
operator matrices and sorted link lists, or the single-excitation closure)
with no CAS array, up to the reference's 63 orbitals
(pyscf/lib/mcscf/fci_gas.h:24-27).
"""

import ctypes
import itertools
import math

import numpy

from . import _lib
from . import addons_gas
from . import direct_spin1 as _ds

GAS_PSPACE_MATVEC_MAX = 64      # fci_gas.py:38
GAS_SPIN_PSPACE_MATVEC_MAX = 400    # fci_gas.py:39
GAS_INIT_NOISE_NORM = 1e-3      # fci_gas.py:40
# sum(gas_orbs) from which spaces are sector-native by default (the CAS
# embedding's restricted fused kernel covers up to 16 orbitals)
NATIVE_MIN_ORBITALS = 17


def _auto_native(norb):
    return sum(norb) >= NATIVE_MIN_ORBITALS


def _pair(eri, norb):
    return _ds.pair_matrix(eri, norb)


def absorb_h1e(h1e, eri, norb, nelec, fac=0.5):
    """Absorbed two-electron Hamiltonian as an (npair, npair) pair matrix
    (fci_gas.py:132-138)."""
    h1e = numpy.ascontiguousarray(h1e, dtype=numpy.float64).reshape(int(norb), int(norb))
    return _ds.absorb_h1e(h1e, _pair(eri, norb), norb, nelec, fac)


_ORDERED = {}


def _sector_major_space(norb, nelec):
    """The CAS string space of sum(norb) orbitals with addresses in
    sector-major order (strings grouped by their occupation of the
    subspaces, fcis_gas_string_order), shared by the GAS spaces of one
    partition.  One subspace: the colex space of the CAS solver."""
    key = (tuple(norb), tuple(nelec))
    if len(norb) == 1:
        return _ds.get_space(norb[0], nelec)
    sp = _ORDERED.get(key)
    if sp is None:
        lib = _lib.load()
        n = sum(norb)
        na = numpy.ascontiguousarray(norb, dtype=numpy.int32)
        strs = []
        for k in nelec:
            out = numpy.empty(math.comb(n, k), dtype=numpy.uint64)
            _lib.check(lib.fcis_gas_string_order(len(norb), na.ctypes.data_as(
                ctypes.POINTER(ctypes.c_int)), int(k), out.ctypes.data_as(
                ctypes.POINTER(ctypes.c_uint64))), "fcis_gas_string_order")
            strs.append(out)
        sp = _ORDERED[key] = _ds.Space(n, nelec, strings=tuple(strs))
    return sp


class GasSpace:
    """Device GAS space for subspace sizes ``norb`` and block set ``blocks``
    (GasSpace, fci_gas.py:420-567).  Usable as a context manager; close()
    releases the device maps.  A CAS-embedded space (``native=False``) runs
    its kernels on a CAS string space of sum(norb) orbitals in sector-major
    address order (``cas``), so that each sector is one address range; a
    sector-native space (``native=True``; the default from
    NATIVE_MIN_ORBITALS orbitals) has no CAS space (``cas`` is None) and
    plans of its own (``make_plan``).  gas2fci / fci2gas use a map to the
    colex (pyscf) CAS layout, built when first needed."""

    def __init__(self, norb, nelec, blocks=None, lib=None, compress_links=False, native=None):
        self.norb = tuple(int(x) for x in norb)
        self.nelec = _ds._unpack_nelec(nelec)
        if blocks is None:
            if len(self.norb) != 1:
                raise ValueError("blocks are required for multi-space GAS")
            blocks = numpy.asarray([self.nelec], dtype=numpy.int32)
        self.blocks = addons_gas.normalize_blocks(blocks, len(self.norb))
        self.limits = addons_gas.check_kernel_limits(self.norb, self.nelec, self.blocks)
        self.lib = lib or _lib.load()
        self.native = _auto_native(self.norb) if native is None else bool(native)
        self._compressed = bool(compress_links)
        if self.native:
            self.cas = None
            self.handle = self._create_native()
        else:
            self.cas = _sector_major_space(self.norb, self.nelec)
            self.handle = self._create(self.cas)
        self._colex_cas_ = None
        self._colex = None
        nd, nb, ns = ctypes.c_int64(), ctypes.c_int(), ctypes.c_int()
        _lib.check(self.lib.fcis_gas_space_info(self.handle, ctypes.byref(nd), ctypes.byref(nb),
                                                ctypes.byref(ns)), "fcis_gas_space_info")
        self._ndet, self._nblock, self._nsector = nd.value, nb.value, ns.value
        self._blocks_desc = None

    def _create(self, cas):
        norb_a = numpy.ascontiguousarray(self.norb, dtype=numpy.int32)
        occ = numpy.ascontiguousarray(self.blocks.reshape(-1), dtype=numpy.int32)
        h = ctypes.c_void_p()
        _lib.check(self.lib.fcis_gas_space_create(
            ctypes.byref(h), cas.handle, len(self.norb),
            norb_a.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), int(self.blocks.shape[0]),
            occ.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _lib.stream_ptr()),
            "fcis_gas_space_create")
        return h

    def _create_native(self):
        norb_a = numpy.ascontiguousarray(self.norb, dtype=numpy.int32)
        occ = numpy.ascontiguousarray(self.blocks.reshape(-1), dtype=numpy.int32)
        h = ctypes.c_void_p()
        ip = ctypes.POINTER(ctypes.c_int)
        _lib.check(self.lib.fcis_gas_space_create_native(
            ctypes.byref(h), len(self.norb), norb_a.ctypes.data_as(ip), int(self.nelec[0]),
            int(self.nelec[1]), int(self.blocks.shape[0]), occ.ctypes.data_as(ip),
            _lib.stream_ptr()), "fcis_gas_space_create_native")
        return h

    @property
    def _colex_cas(self):
        """The colex CAS space of sum(norb) orbitals (gas2fci / fci2gas)."""
        if self._colex_cas_ is None:
            self._colex_cas_ = _ds.get_space(sum(self.norb), self.nelec)
        return self._colex_cas_

    def _colex_handle(self):
        """GAS space on the colex CAS space (for gas2fci / fci2gas)."""
        self._live()
        if self._colex is None:
            if self.cas is not None and self._colex_cas is self.cas:
                self._colex = self.handle
            else:
                self._colex = self._create(self._colex_cas)
        return self._colex

    def native_info(self):
        """(determinants within one excitation of the space, the longest
        closure row, alpha strings, beta strings) of a sector-native space's
        tables; zeros otherwise."""
        nat, n1, mr = ctypes.c_int(), ctypes.c_int64(), ctypes.c_int64()
        na, nb = ctypes.c_int64(), ctypes.c_int64()
        _lib.check(self.lib.fcis_gas_native_info(self._live(), ctypes.byref(nat), ctypes.byref(n1),
                                                 ctypes.byref(mr), ctypes.byref(na),
                                                 ctypes.byref(nb)), "fcis_gas_native_info")
        return n1.value, mr.value, na.value, nb.value

    def make_plan(self, h2e, precision="fp64", workspace_bytes=0):
        """Sigma plan of a sector-native space (the Davidson driver's hook);
        None for a CAS-embedded space, whose solves run SigmaPlans on
        ``cas`` through ``hop``."""
        if not self.native:
            return None
        return GasContractPlan(self, h2e, precision, workspace_bytes)

    def work_bytes(self, precision="fp64"):
        """Device bytes a Davidson solve needs beside its vectors and plans:
        the CAS embedding buffers (none for a sector-native space)."""
        if self.native:
            return 0
        return (2 if precision == "fp64" else 4) * self.cas.ndet * 8

    def det_strings(self, addr):
        """(alpha strings, beta strings) of the GAS determinants ``addr``."""
        a = numpy.ascontiguousarray(addr, dtype=numpy.int64).reshape(-1)
        sa = numpy.empty(a.size, dtype=numpy.uint64)
        sb = numpy.empty(a.size, dtype=numpy.uint64)
        _lib.check(self.lib.fcis_gas_det_strings(
            self._live(), a.size, a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
            sa.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
            sb.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))), "fcis_gas_det_strings")
        return sa, sb

    def embed_colex(self, gas_vec, cas_out, nvec=1):
        """cas_out[v] <- gas2fci(gas_vec[v]) in pyscf's (colex) CAS layout."""
        _lib.check(self.lib.fcis_gas_embed(self._colex_handle(), _lib.vptr(gas_vec),
                                           _lib.vptr(cas_out), int(nvec), _lib.stream_ptr()),
                   "fcis_gas_embed")

    def project_colex(self, cas_vec, gas_out, nvec=1):
        """gas_out[v] <- fci2gas(cas_vec[v]) from pyscf's (colex) CAS layout."""
        _lib.check(self.lib.fcis_gas_project(self._colex_handle(), _lib.vptr(cas_vec),
                                             _lib.vptr(gas_out), int(nvec), _lib.stream_ptr()),
                   "fcis_gas_project")

    # -- descriptors (fci_gas.py:459-542) -------------------------------
    def _live(self):
        if self.handle is None:
            raise RuntimeError("GAS space is closed")
        return self.handle

    @property
    def ndet(self):
        return self._ndet

    @property
    def n(self):
        """Vector length of this basis (the Davidson driver's basis size)."""
        return self._ndet

    @property
    def nsector(self):
        return self._nsector

    @property
    def nblock(self):
        return self._nblock

    @property
    def norb_total(self):
        return sum(self.norb)

    def core_info(self):
        self._live()
        start = tuple(int(x) for x in numpy.cumsum((0,) + self.norb)[:-1])
        return {"ngas": len(self.norb), "norb": self.norb, "start": start,
                "norb_total": self.norb_total, "nelec": self.nelec, "ndet": self.ndet,
                "nsector": self.nsector, "nblock": self.nblock}

    def compress_links(self):
        """The device link tables serve sigma, diagonal and RDMs alike;
        kept for interface parity (gas_space_compress_links)."""
        self._compressed = True

    def links_are_compressed(self):
        return self._compressed

    def addr2str(self, sector, addr):
        s = ctypes.c_uint64()
        _lib.check(self.lib.fcis_gas_addr2str(self._live(), int(sector), int(addr),
                                              ctypes.byref(s)), "fcis_gas_addr2str")
        return int(s.value)

    def str2addr(self, sector, string):
        a = ctypes.c_int64()
        _lib.check(self.lib.fcis_gas_str2addr(self._live(), int(sector), int(string),
                                              ctypes.byref(a)), "fcis_gas_str2addr")
        return int(a.value)

    def sector_occupations(self):
        occ = numpy.zeros((self.nsector, len(self.norb)), dtype=numpy.int32)
        nstr = numpy.zeros(self.nsector, dtype=numpy.int64)
        _lib.check(self.lib.fcis_gas_sector_info(
            self._live(), occ.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
            nstr.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))), "fcis_gas_sector_info")
        return occ, nstr

    def block_descriptors(self):
        if self._blocks_desc is None:
            n = self.nblock
            off, na, nb = (numpy.zeros(n, dtype=numpy.int64) for _ in range(3))
            sa, sb = numpy.zeros(n, dtype=numpy.int32), numpy.zeros(n, dtype=numpy.int32)
            ip, lp = ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int64)
            _lib.check(self.lib.fcis_gas_block_info(
                self._live(), off.ctypes.data_as(lp), sa.ctypes.data_as(ip),
                sb.ctypes.data_as(ip), na.ctypes.data_as(lp), nb.ctypes.data_as(lp)),
                "fcis_gas_block_info")
            self._blocks_desc = [dict(offset=int(a), sa=int(b), sb=int(c), na=int(d), nb=int(e))
                                 for a, b, c, d, e in zip(off, sa, sb, na, nb)]
        for d in self._blocks_desc:
            yield dict(d)

    def block_ndet(self, block_id):
        d = list(self.block_descriptors())[int(block_id)]
        return d["na"] * d["nb"]

    # -- device transforms ----------------------------------------------
    def embed_dev(self, gas_vec, cas_out, nvec=1):
        """cas_out[v] <- gas_vec[v] embedded in the kernels' CAS layout
        (``self.cas`` address order; device float64, contiguous)."""
        _lib.check(self.lib.fcis_gas_embed(self._live(), _lib.vptr(gas_vec), _lib.vptr(cas_out),
                                           int(nvec), _lib.stream_ptr()), "fcis_gas_embed")

    def project_dev(self, cas_vec, gas_out, nvec=1):
        """gas_out[v] <- cas_vec[v] (``self.cas`` layout) projected on the space."""
        _lib.check(self.lib.fcis_gas_project(self._live(), _lib.vptr(cas_vec),
                                             _lib.vptr(gas_out), int(nvec), _lib.stream_ptr()),
                   "fcis_gas_project")

    def sigma_dev(self, plan, x, out, work):
        """out <- P H P x with a CAS SigmaPlan on ``self.cas`` (any
        precision); work: 2 * cas.ndet float64 device scratch."""
        if x.data_ptr() == out.data_ptr():
            raise ValueError("sigma input and output must not alias")
        _lib.check(self.lib.fcis_gas_sigma_execute(self._live(), plan.handle, _lib.vptr(x),
                                                   _lib.vptr(out), _lib.vptr(work),
                                                   _lib.stream_ptr()), "fcis_gas_sigma_execute")
        return out

    def hop(self, plan):
        """x (GAS) -> P H P x through a CAS sigma plan on ``self.cas`` (a
        sector-native space's own plan applies directly)."""
        import torch
        if self.native:
            return plan
        work = torch.empty(2 * self.cas.ndet, dtype=torch.float64, device="cuda")

        def apply(x, out):
            self.sigma_dev(plan, x, out, work)
        return apply

    def rows_info(self):
        """(alpha rows, beta chunks, CAS chunks) of the restricted fused
        sigma: the fraction of the CAS sigma's work it does."""
        a, b, c = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        _lib.check(self.lib.fcis_gas_rows_info(self._live(), ctypes.byref(a), ctypes.byref(b),
                                               ctypes.byref(c)), "fcis_gas_rows_info")
        return a.value, b.value, c.value

    def close(self):
        if getattr(self, "handle", None) is not None:
            if self._colex is not None and self._colex is not self.handle:
                self.lib.fcis_gas_space_free(self._colex)
            self._colex = None
            self.lib.fcis_gas_space_free(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_GAS_SPACES = {}


def _givens_factors(w):
    """Adjacent-row Givens rotations (p, p + 1, c, s) that take the square
    orthogonal w to a diagonal D = +-1, column by column from the bottom
    (rows p, q <- c row_p + s row_q, -s row_p + c row_q), in the order
    made, and diag(D)."""
    w = numpy.array(w, dtype=numpy.float64)
    n = w.shape[0]
    rots = []
    for col in range(n - 1):
        for q in range(n - 1, col, -1):
            p = q - 1
            a, b = w[p, col], w[q, col]
            if b == 0.0:
                continue
            r = math.hypot(a, b)
            c, s = a / r, b / r
            w[[p, q]] = numpy.array([[c, s], [-s, c]]) @ w[[p, q]]
            rots.append((p, q, c, s))
    return rots, numpy.where(numpy.diag(w) < 0, -1.0, 1.0)


def _cached_space(norb, nelec, blocks, native=None):
    native = _auto_native(tuple(norb)) if native is None else bool(native)
    key = (tuple(norb), tuple(nelec), numpy.asarray(blocks, dtype=numpy.int32).tobytes(), native)
    g = _GAS_SPACES.get(key)
    if g is None:
        g = _GAS_SPACES[key] = GasSpace(norb, nelec, blocks, native=native)
    return g


def _vector_in(x, n):
    """(device tensor of n float64, was-host flag, original shape)."""
    import torch
    if isinstance(x, torch.Tensor):
        t = x.detach().to(device="cuda", dtype=torch.float64).contiguous().reshape(-1)
        host, shape = False, tuple(x.shape)
    else:
        a = numpy.ascontiguousarray(x, dtype=numpy.float64)
        shape = a.shape
        t = torch.from_numpy(a.reshape(-1)).to("cuda")
        host = True
    if t.numel() != n:
        raise ValueError("CI vector size does not match GAS determinant count")
    return t, host, shape


def gas2fci(fcivec, gas):
    """GAS vector -> CAS (nstr_alpha, nstr_beta) array, zeros outside the
    GAS space (fci_gas.py:141-182).  numpy in, numpy out; device tensors
    stay on the device."""
    import torch
    c, host, _ = _vector_in(fcivec, gas.ndet)
    out = torch.empty(gas._colex_cas.shape, dtype=torch.float64, device="cuda")
    gas.embed_colex(c, out)
    return out.cpu().numpy() if host else out


def fci2gas(fcivec, gas):
    """CAS array -> GAS vector in canonical block order (fci_gas.py:185-208)."""
    import torch
    if isinstance(fcivec, torch.Tensor):
        c, host = fcivec.detach().to(device="cuda", dtype=torch.float64).contiguous(), False
    else:
        c, host = torch.from_numpy(numpy.ascontiguousarray(fcivec, dtype=numpy.float64)).cuda(), True
    if c.numel() != gas._colex_cas.ndet:
        raise ValueError("CI vector size does not match full FCI determinant count")
    out = torch.empty(gas.ndet, dtype=torch.float64, device="cuda")
    gas.project_colex(c.reshape(-1), out)
    return out.cpu().numpy() if host else out


class GasContractPlan:
    """Reusable restricted sigma plan (GasContractPlan, fci_gas.py:570-650),
    through fcis_gas_sigma_plan_*.  The space must outlive the plan."""

    def __init__(self, gas, eri, precision="fp64", workspace_bytes=0):
        if gas.handle is None:
            raise RuntimeError("GAS space is closed")
        self.lib = gas.lib
        self.gas = gas
        self.eri = _pair(eri, gas.norb_total)
        prec = _lib.PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        self.precision = prec
        h = ctypes.c_void_p()
        _lib.check(self.lib.fcis_gas_sigma_plan_create(
            ctypes.byref(h), gas.handle, _lib.dptr(self.eri), prec,
            int(workspace_bytes) or _ds._default_workspace()), "fcis_gas_sigma_plan_create")
        self._plan = h

    @property
    def computes_fp64(self):
        """True when the product is FP64 whatever the requested precision
        (an FP64 plan, or the sector-native direct forms)."""
        return self.precision == _lib.PRECISIONS["fp64"] or self.path in ("native-direct", "native-links")

    @property
    def ndet(self):
        return self.gas.ndet

    @property
    def norb(self):
        return self.gas.norb_total

    @property
    def nelec(self):
        return self.gas.nelec

    def __call__(self, c, out):
        """out <- P H P c (the Davidson driver's operator)."""
        return self.execute(c, out)

    @property
    def path(self):
        """"cas-three-kernel", "cas-fused", "native-t1" (closure GEMM),
        "native-direct" (operator form) or "native-links" (link-pair loops)."""
        if self._plan is None:
            raise RuntimeError("Hamiltonian contraction plan is closed")
        k = self.lib.fcis_gas_sigma_plan_path(self._plan)
        return {0: "cas-three-kernel", 1: "cas-fused", 2: "native-t1", 3: "native-direct",
                4: "native-links"}[k]

    def execute(self, c, out):
        """out <- P H P c for device tensors."""
        if self._plan is None or self.gas is None:
            raise RuntimeError("Hamiltonian contraction plan is closed")
        if c.data_ptr() == out.data_ptr():
            raise ValueError("sigma input and output must not alias")
        _lib.check(self.lib.fcis_gas_sigma_plan_execute(self._plan, _lib.vptr(c),
                                                        _lib.vptr(out), _lib.stream_ptr()),
                   "fcis_gas_sigma_plan_execute")
        return out

    def contract(self, fcivec):
        """P H P c; same shape and residence (host/device) as the input."""
        import torch
        if self._plan is None or self.gas is None:
            raise RuntimeError("Hamiltonian contraction plan is closed")
        c, host, shape = _vector_in(fcivec, self.ndet)
        out = torch.empty_like(c)
        self.execute(c, out)
        return out.cpu().numpy().reshape(shape) if host else out.reshape(shape)

    def close(self):
        if getattr(self, "_plan", None) is not None:
            self.lib.fcis_gas_sigma_plan_free(self._plan)
            self._plan = None
        self.gas = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class GasSpinPlan:
    """S^2 on GAS vectors (_GasSpinPlan, fci_gas.py:266-419): the CAS S^2
    kernel between embedding and projection, or on a sector-native space
    S_z(S_z+1) + N_beta - sum_pq E^a_pq E^b_qp on the space's own tables;
    exact when the block set is spin complete (checked by the callers, as in
    the reference).  Holds the space, which must stay open."""

    def __init__(self, gas):
        import torch
        self.gas = gas
        self.ndet = gas.ndet
        na, nb = gas.nelec
        sz = 0.5 * (na - nb)
        self.diagonal = sz * (sz + 1.0) + nb
        if gas.cas is None:
            # sector-native: the library applies S^2 on the space's own tables
            self._plan = self._work = None
        else:
            self._plan = _ds.SpinPlan(gas.cas)
            self._work = torch.empty(2 * gas.cas.ndet, dtype=torch.float64, device="cuda")

    def contract_dev(self, x, out):
        """out <- S^2 x for device GAS vectors."""
        if x.data_ptr() == out.data_ptr():
            raise ValueError("S^2 input and output must not alias")
        _lib.check(self.gas.lib.fcis_gas_contract_ss(
            self.gas._live(), None if self._plan is None else self._plan.handle, _lib.vptr(x),
            _lib.vptr(out), _lib.vptr(self._work), _lib.stream_ptr()), "fcis_gas_contract_ss")
        return out

    def contract(self, fcivec):
        import torch
        c, host, shape = _vector_in(fcivec, self.ndet)
        out = self.contract_dev(c, torch.empty_like(c))
        return out.cpu().numpy().reshape(shape) if host else out.reshape(shape)

    def penalty(self, vector, penalty):
        """P x for a SpinPenalty P, both S^2 applications on the device;
        numpy in, numpy out (device tensors stay on the device)."""
        import torch
        x, host, shape = _vector_in(vector, self.ndet)
        y = penalty.apply(lambda v: self.contract_dev(v, torch.empty_like(v)), x)
        return y.cpu().numpy().reshape(shape) if host else y.reshape(shape)

    def diagonal_dev(self):
        import torch
        if self._plan is None:
            out = torch.empty(self.ndet, dtype=torch.float64, device="cuda")
            _lib.check(self.gas.lib.fcis_gas_ss_diag(self.gas._live(), _lib.vptr(out),
                                                     _lib.stream_ptr()), "fcis_gas_ss_diag")
            return out
        full = torch.empty(self.gas.cas.ndet, dtype=torch.float64, device="cuda")
        self._plan.diag(full)
        out = torch.empty(self.ndet, dtype=torch.float64, device="cuda")
        self.gas.project_dev(full, out)
        return out

    def diagonal_vector(self):
        """Exact determinant-basis diagonal of S^2 (fci_gas.py:374-398)."""
        return self.diagonal_dev().cpu().numpy()

    def matrix(self):
        """Dense S^2 of a bounded space (fci_gas.py:400-409)."""
        import torch
        x = torch.zeros(self.ndet, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        M = numpy.empty((self.ndet, self.ndet))
        for j in range(self.ndet):
            x.zero_()
            x[j] = 1.0
            M[:, j] = self.contract_dev(x, y).cpu().numpy()
        return 0.5 * (M + M.T)

    def project(self, vector, target, eigenvalues):
        """Projection on the S(S+1) = target eigenspace (fci_gas.py:411-419):
        prod over the other allowed eigenvalues l of (S^2 - l) / (target - l)."""
        import torch
        from .davidson import DeviceOps
        ops = DeviceOps()
        x, host, shape = _vector_in(vector, self.ndet)
        x = x.clone()
        t = torch.empty_like(x)
        for lam in eigenvalues:
            if abs(lam - target) < 1e-10:
                continue
            self.contract_dev(x, t)
            ops.lincomb(1, self.ndet, x, numpy.array([-lam]), 1.0, t)
            ops.lincomb(1, self.ndet, t, numpy.array([1.0 / (target - lam)]), 0.0, x)
        return x.cpu().numpy().reshape(shape) if host else x.reshape(shape)


class SpinPenalty:
    """The GAS solver's spin penalty P = shift * (S^2 - t), or
    shift * (S^2 - t)^2 when t lies above the lowest S(S+1) the electron
    counts allow (the linear form would reward spins below t).  The
    penalty the reference's fixed-orbital solver and Newton driver add
    (fci_gas.py:211-263).

    With m = |N_alpha - N_beta| / 2 and at most min(N, 2 norb - N) unpaired
    electrons, S runs over m, m + 1, ..., min(N, 2 norb - N) / 2, and t must
    be one of those S(S+1) (default: the lowest, m(m+1))."""

    __slots__ = ("shift", "target", "spectrum", "quadratic")

    def __init__(self, shift, target, spectrum):
        self.shift = float(shift)
        self.target = float(target)
        self.spectrum = tuple(float(x) for x in spectrum)
        self.quadratic = self.target > self.spectrum[0] + 1e-10

    @classmethod
    def of(cls, solver, norb, nelec):
        """The penalty a solver carries (``ss_penalty`` / ``ss_value``, as set
        by fix_spin_), or None."""
        shift = getattr(solver, "ss_penalty", None)
        if shift is None:
            return None
        shift = float(shift)
        if not (numpy.isfinite(shift) and shift >= 0.0):
            raise ValueError("spin penalty shift must be finite and nonnegative")
        na, nb = _ds._unpack_nelec(nelec)
        unpaired_max = min(na + nb, 2 * int(norb) - na - nb)
        s_values = 0.5 * numpy.arange(abs(na - nb), unpaired_max + 1, 2)
        spectrum = s_values * (s_values + 1.0)
        want = getattr(solver, "ss_value", None)
        target = spectrum[0] if want is None else float(want)
        if not numpy.isfinite(target):
            raise ValueError("target S^2 must be finite")
        if target < spectrum[0] - 1e-10:
            raise ValueError("target S^2 is below the minimum allowed by Nalpha-Nbeta")
        if numpy.min(numpy.abs(spectrum - target)) >= 1e-10:
            raise ValueError("target S^2 is not an allowed S(S+1) value for this space")
        return cls(shift, target, spectrum)

    def active(self):
        return self.shift != 0.0

    def apply(self, s2, x):
        """P x for a vector x and ``s2`` mapping a vector to S^2 times it
        (arrays in, array out; x is not modified)."""
        y = s2(x) - self.target * x
        if self.quadratic:
            y = s2(y) - self.target * y
        return self.shift * y

    def diagonal(self, s2_diagonal):
        """Preconditioner diagonal: shift * (diag(S^2) - t), squared
        entrywise in the quadratic case (an approximation of the diagonal of
        (S^2 - t)^2, whose exact value needs the off-diagonal S^2 row)."""
        d = s2_diagonal - self.target
        return self.shift * (d * d if self.quadratic else d)

    def value(self, x, s2x):
        """<x|P|x> for a normalised x with s2x = S^2 x (device or host)."""
        if self.quadratic:
            d = s2x - self.target * x
            return self.shift * float((d * d).sum())
        return self.shift * (float((x * s2x).sum()) - self.target)


def fix_spin_(fciobj, shift=0.2, ss=None):
    """Attach a spin penalty shift * (S^2 - ss) (or its square when ss is
    above the minimum) to a GAS solver; ss None = the lowest S(S+1) allowed
    by Nalpha - Nbeta (the attributes the reference's GAS kernel reads,
    fci_gas.py:211-239)."""
    fciobj.ss_penalty = shift
    fciobj.ss_value = ss
    return fciobj


class _PenaltyBasis:
    """GAS basis whose H application adds the spin penalty P (SpinPenalty)."""

    def __init__(self, gas, spin, params):
        self.gas, self.spin, self.params = gas, spin, params
        self.n = gas.ndet

    def penalty(self, x, out, t1, t2, ops):
        """out += P x with two scratch vectors (no allocation per call)."""
        pen, n = self.params, self.n
        self.spin.contract_dev(x, t1)
        ops.lincomb(1, n, x, numpy.array([-pen.target]), 1.0, t1)
        if pen.quadratic:
            self.spin.contract_dev(t1, t2)
            ops.lincomb(1, n, t1, numpy.array([-pen.target]), 1.0, t2)
            t1 = t2
        ops.lincomb(1, n, t1, numpy.array([pen.shift]), 1.0, out)

    def make_plan(self, h2e, precision="fp64", workspace_bytes=0):
        return self.gas.make_plan(h2e, precision, workspace_bytes)

    def work_bytes(self, precision="fp64"):
        return self.gas.work_bytes(precision) + 2 * self.n * 8

    def hop(self, plan):
        import torch
        from .davidson import DeviceOps
        base = self.gas.hop(plan)
        ops = DeviceOps()
        t1 = torch.empty(self.n, dtype=torch.float64, device="cuda")
        t2 = torch.empty_like(t1)

        def apply(x, out):
            base(x, out)
            self.penalty(x, out, t1, t2, ops)
        return apply


class GasRDMPlan:
    """RDMs of GAS vectors (GasRDMPlan, fci_gas.py:653-757)."""

    def __init__(self, solver, norb, nelec, gas=None):
        self.lib = _lib.load()
        self._owned = gas is None
        self.gas = solver.make_space(norb, nelec) if gas is None else gas

    @property
    def ndet(self):
        return self.gas.ndet

    @property
    def norb(self):
        return self.gas.norb_total

    def _rdm(self, bra, ket, with_dm2):
        if self.gas is None or self.gas.handle is None:
            raise RuntimeError("RDM plan is closed")
        b, _, _ = _vector_in(bra, self.ndet)
        k, _, _ = _vector_in(ket, self.ndet)
        n = self.norb
        d1a, d1b = numpy.zeros((n, n)), numpy.zeros((n, n))
        d2 = [numpy.zeros((n,) * 4) for _ in range(3)] if with_dm2 else [None] * 3
        _lib.check(self.lib.fcis_gas_trans_rdm12s(
            self.gas.handle, _lib.vptr(b), _lib.vptr(k), _lib.dptr(d1a), _lib.dptr(d1b),
            *[None if x is None else _lib.dptr(x) for x in d2], _lib.stream_ptr()),
            "fcis_gas_trans_rdm12s")
        return (d1a, d1b), tuple(d2)

    def make_rdm1s(self, cibra, ciket):
        return self._rdm(cibra, ciket, False)[0]

    def make_rdm1(self, cibra, ciket):
        a, b = self.make_rdm1s(cibra, ciket)
        return a + b

    def make_rdm12s(self, cibra, ciket):
        return self._rdm(cibra, ciket, True)

    def make_rdm12(self, cibra, ciket):
        (a, b), (aa, ab, bb) = self.make_rdm12s(cibra, ciket)
        return a + b, aa + bb + ab + ab.transpose(2, 3, 0, 1)

    def close(self):
        """Release the plan (and its space when the plan created it)."""
        if getattr(self, "gas", None) is not None:
            if self._owned:
                self.gas.close()
            self.gas = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class FCISolver(_ds.FCISolver):
    """Determinant-space GAS solver (fci_gas.FCISolver, fci_gas.py:760-1483)
    with every O(ndet) step on the GPU.  Restrictions come from
    ``gas_orbs``, ``gas_restr`` and ``gas_restr_type`` (the four forms of
    addons_gas.normalize_gas_spec); CI vectors are 1-D in GAS block order.
    ``precision`` selects the Davidson sigma arithmetic as in the CAS
    solver (FP64 refinement for "fp32" / "bf16x3"; a sector-native space
    computes in fp64 or fp32).  ``native`` selects the space kind: None
    (default) embeds spaces of up to 16 orbitals in the CAS space and makes
    larger ones sector-native, True / False force one kind."""

    def __init__(self, mol=None, gas_orbs=None, gas_restr=None,
                 gas_restr_type=addons_gas.GAS_RESTR_SPIN_SUPERGROUP, lib=None):
        super().__init__(mol)
        self.gas_orbs = tuple(gas_orbs) if gas_orbs is not None else None
        self.gas_restr = gas_restr
        self.gas_restr_type = gas_restr_type
        self.lib = lib or _lib.load()
        self.e_spin_penalty = None
        self.e_physical = None
        self.spin_penalty_method = None
        self.native = None

    # -- spaces ----------------------------------------------------------
    def _space_spec(self, norb, nelec):
        nelec = _ds._unpack_nelec(nelec, self.spin)
        if self.gas_orbs is None:
            orbs, D = addons_gas.normalize_gas_spec((int(norb),), nelec)
        else:
            if sum(self.gas_orbs) != int(norb):
                raise ValueError("sum(gas_orbs) must equal norb")
            orbs, D = addons_gas.normalize_gas_spec(self.gas_orbs, nelec, self.gas_restr,
                                                    self.gas_restr_type)
        return orbs, nelec, D

    def space_info(self, norb, nelec):
        orbs, nelec, D = self._space_spec(norb, nelec)
        return addons_gas.check_kernel_limits(orbs, nelec, D)

    def make_space(self, norb, nelec, compress_links=False):
        """An owned GasSpace for this solver's restriction."""
        orbs, nelec, D = self._space_spec(norb, nelec)
        return GasSpace(orbs, nelec, D, lib=self.lib, compress_links=compress_links,
                        native=getattr(self, "native", None))

    def _space(self, norb, nelec):
        orbs, nelec, D = self._space_spec(norb, nelec)
        return _cached_space(orbs, nelec, D, getattr(self, "native", None))

    def make_rdm_plan(self, norb, nelec):
        return GasRDMPlan(self, norb, nelec)

    def make_contract_plan(self, eri, norb, nelec, precision="fp64"):
        return GasContractPlan(self._space(norb, nelec), eri, precision)

    def make_spin_plan(self, norb, nelec):
        """S^2 plan on this restriction; requires a spin-complete block set
        (fci_gas.py:853-869)."""
        orbs, nelec, D = self._space_spec(norb, nelec)
        if not addons_gas.is_spin_complete(orbs, nelec, D):
            raise ValueError("make_spin_plan requires a spin-complete GAS restriction")
        return GasSpinPlan(self._space(norb, nelec))

    def contract_ss(self, fcivec, norb, nelec):
        """S^2 c on a GAS vector (fci_gas.py:993-1001)."""
        orbs, nelec, D = self._space_spec(norb, nelec)
        if not addons_gas.is_spin_complete(orbs, nelec, D):
            raise ValueError("contract_ss requires a spin-complete GAS restriction")
        return GasSpinPlan(self._space(norb, nelec)).contract(fcivec)

    def _check_plan(self, plan, norb, nelec, kind, cls):
        if not isinstance(plan, cls):
            raise TypeError("plan must be a %s" % cls.__name__)
        if plan.gas is None or plan.gas.handle is None:
            raise RuntimeError(kind + " plan is closed")
        orbs, nelec, D = self._space_spec(norb, nelec)
        if (tuple(orbs) != plan.gas.norb or tuple(nelec) != tuple(plan.gas.nelec)
                or not numpy.array_equal(D, plan.gas.blocks)):
            raise ValueError(kind + " plan does not match the GAS space")

    # -- pyscf protocol ----------------------------------------------------
    def get_init_guess(self, norb, nelec, nroots, hdiag, gas=None):
        """Lowest-diagonal unit vectors plus global noise, GAS layout
        (fci_gas.py:881-902)."""
        n = (self.space_info(norb, nelec)["ndet_estimate"] if gas is None else gas.ndet)
        hd = numpy.asarray(hdiag).reshape(-1) if not hasattr(hdiag, "device") else hdiag
        if hd.shape[0] != n:
            raise ValueError("hdiag size does not match GAS determinant count")
        if not 1 <= int(nroots) <= n:
            raise ValueError("nroots must be between 1 and the GAS dimension")
        return [g.cpu().numpy() for g in self._init_guess_dev(hd, int(nroots))]

    def make_hdiag(self, h1e, eri, norb, nelec, *args, **kwargs):
        return self._hdiag_dev(h1e, eri, norb, nelec).cpu().numpy()

    def _hdiag_dev(self, h1e, eri, norb, nelec, gas=None):
        import torch
        gas = gas or self._space(norb, nelec)
        n = gas.norb_total
        h1e = numpy.ascontiguousarray(h1e, dtype=numpy.float64).reshape(n, n)
        eri = _pair(eri, n)
        out = torch.empty(gas.ndet, dtype=torch.float64, device="cuda")
        _lib.check(self.lib.fcis_gas_make_hdiag(gas.handle, _lib.dptr(h1e), _lib.dptr(eri),
                                                _lib.vptr(out), _lib.stream_ptr()),
                   "fcis_gas_make_hdiag")
        return out

    def contract_2e(self, eri, fcivec, norb, nelec, link_index=None, *args, **kwargs):
        """P H P c for the absorbed pair matrix ``eri``; with ``plan=`` (a
        GasContractPlan) the plan's Hamiltonian is used and ``eri`` ignored."""
        plan = kwargs.pop("plan", None)
        if plan is not None:
            self._check_plan(plan, norb, nelec, "contraction", GasContractPlan)
            return plan.contract(fcivec)
        with GasContractPlan(self._space(norb, nelec), eri,
                             kwargs.pop("precision", "fp64")) as p:
            return p.contract(fcivec)

    def contract_1e(self, f1e, fcivec, norb, nelec, link_index=None, **kwargs):
        raise NotImplementedError(
            "GASCI uses the absorbed one-electron Hamiltonian and contract_2e")

    def energy(self, h1e, eri, fcivec, norb, nelec, link_index=None):
        import torch
        h2e = absorb_h1e(h1e, eri, norb, _ds._unpack_nelec(nelec, self.spin), 0.5)
        gas = self._space(norb, nelec)
        c, _, _ = _vector_in(fcivec, gas.ndet)
        with GasContractPlan(gas, h2e) as p:
            hc = p.execute(c, torch.empty_like(c))
        return float(torch.dot(c, hc))

    def pspace(self, h1e, eri, norb, nelec, hdiag=None, np=400):
        """(addresses, H block) over the lowest-diagonal determinants,
        at most GAS_PSPACE_MATVEC_MAX of them (fci_gas.py:955-988)."""
        import torch
        if int(np) < 1:
            raise ValueError("np must be positive")
        gas = self._space(norb, nelec)
        nelec = _ds._unpack_nelec(nelec, self.spin)
        hd = self._hdiag_dev(h1e, eri, norb, nelec, gas) if hdiag is None else \
            torch.as_tensor(numpy.asarray(hdiag, dtype=numpy.float64).reshape(-1)).cuda()
        if hd.numel() != gas.ndet:
            raise ValueError("hdiag size does not match GAS determinant count")
        m = min(int(np), gas.ndet, GAS_PSPACE_MATVEC_MAX)
        addr = torch.topk(hd, m, largest=False).indices
        return addr.cpu().numpy(), self._h_block(gas, h1e, eri, addr)

    def _h_block(self, gas, h1e, eri, addr):
        """H over the GAS determinants ``addr``: the Slater-Condon block of
        their strings (fcis_gas_pspace, either kind of space)."""
        import torch
        a = numpy.ascontiguousarray(addr.cpu().numpy() if hasattr(addr, "cpu") else addr,
                                    dtype=numpy.int64).reshape(-1)
        n = gas.norb_total
        h1e = numpy.ascontiguousarray(h1e, dtype=numpy.float64).reshape(n, n)
        eri = _pair(eri, n)
        out = torch.empty((a.size, a.size), dtype=torch.float64, device="cuda")
        _lib.check(self.lib.fcis_gas_pspace(gas._live(), _lib.dptr(h1e), _lib.dptr(eri), a.size,
                                            a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                            _lib.vptr(out), _lib.stream_ptr()), "fcis_gas_pspace")
        return out.cpu().numpy()

    def gen_linkstr(self, norb, nelec, tril=True, spin=None):
        raise NotImplementedError(
            "GAS link tables are owned by GasSpace and are not CAS linkstr arrays")

    def large_ci(self, fcivec, norb, nelec, tol=0.1, return_strs=True):
        """(coefficient, alpha string, beta string) of the large GAS
        coefficients (fci_gas.py:1093-1128)."""
        gas = self._space(norb, nelec)
        ci = numpy.asarray(fcivec.cpu() if hasattr(fcivec, "cpu") else fcivec,
                           dtype=numpy.float64).reshape(-1)
        if ci.size != gas.ndet:
            raise ValueError("CI vector size does not match GAS determinant count")
        sel = numpy.flatnonzero(numpy.abs(ci) > float(tol))
        if sel.size == 0:
            sel = numpy.asarray([int(numpy.argmax(numpy.abs(ci)))])
        blocks = list(gas.block_descriptors())
        offs = numpy.asarray([b["offset"] for b in blocks])
        out = []
        for a in sel.tolist():
            d = blocks[int(numpy.searchsorted(offs, a, side="right")) - 1]
            ia, ib = divmod(a - d["offset"], d["nb"])
            sa, sb = gas.addr2str(d["sa"], ia), gas.addr2str(d["sb"], ib)
            if return_strs:
                out.append((ci[a], bin(sa), bin(sb)))
            else:
                occ = lambda s: numpy.asarray([p for p in range(int(norb)) if (s >> p) & 1],
                                              dtype=numpy.int32)
                out.append((ci[a], occ(sa), occ(sb)))
        return out

    def kernel(self, h1e, eri, norb, nelec, ci0=None, ecore=0, nroots=None, **kwargs):
        """(e + ecore, ci) of the restricted problem (fci_gas.py:1150-1399):
        exact small-space diagonalisation up to GAS_PSPACE_MATVEC_MAX
        determinants, otherwise the device Davidson on P H P."""
        import torch
        nroots = int(nroots if nroots is not None else kwargs.pop("nroots", self.nroots))
        if nroots < 1:
            raise ValueError("nroots must be positive")
        if kwargs.get("orbsym") is not None or kwargs.get("wfnsym") is not None:
            raise NotImplementedError(
                "orbital and wavefunction symmetry filtering is not implemented")

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
        if pspace_size < 0:
            raise ValueError("pspace_size must be nonnegative")
        nelec = _ds._unpack_nelec(nelec, self.spin)
        self.norb, self.nelec = int(norb), nelec
        penalty = SpinPenalty.of(self, norb, nelec)
        self.e_spin_penalty = self.e_physical = self.spin_penalty_method = None
        gas = self._space(norb, nelec)
        if nroots > gas.ndet:
            raise ValueError("nroots exceeds GAS determinant count")
        spin = None
        if penalty is not None:
            if not addons_gas.is_spin_complete(gas.norb, nelec, gas.blocks):
                raise ValueError("fix_spin_ requires a spin-complete GAS restriction")
            spin = GasSpinPlan(gas)
        n = gas.norb_total
        h1e = numpy.ascontiguousarray(h1e, dtype=numpy.float64).reshape(n, n)
        eri = _pair(eri, n)
        h2e = absorb_h1e(h1e, eri, norb, nelec, 0.5)
        hdiag = self._hdiag_dev(h1e, eri, norb, nelec, gas)

        full_pspace = (penalty is None and ci0 is None and not davidson_only
                       and gas.ndet <= pspace_size and gas.ndet <= GAS_PSPACE_MATVEC_MAX)
        full_spin_pspace = (penalty is not None and ci0 is None and pspace_size > 0
                            and gas.ndet <= pspace_size
                            and gas.ndet <= GAS_SPIN_PSPACE_MATVEC_MAX)
        if full_pspace or full_spin_pspace:
            addr = torch.arange(gas.ndet, device="cuda")
            H = self._h_block(gas, h1e, eri, addr)
            if full_spin_pspace:
                S2 = spin.matrix()
                H = H + penalty.apply(lambda v: S2 @ v, numpy.eye(gas.ndet))
                self.spin_penalty_method = "exact-small-space"
            w, v = numpy.linalg.eigh(H)
            e = w[:nroots]
            c = [torch.from_numpy(numpy.ascontiguousarray(v[:, k])).cuda() for k in range(nroots)]
            conv = numpy.ones(nroots, dtype=bool)
        else:
            basis, pre_diag = gas, hdiag
            if penalty is not None:
                ci0 = self._penalty_guesses(gas, spin, penalty, hdiag, ci0, nroots, lindep)
                basis = _PenaltyBasis(gas, spin, penalty)
                pre_diag = hdiag + penalty.diagonal(spin.diagonal_dev())
                self.spin_penalty_method = "projected-plus-global-davidson"
            elif ci0 is not None:
                ci0 = self._guess_list(ci0, nroots, gas.ndet)
            from .mixed import solve
            if gas.native and precision == "bf16x3":
                raise ValueError("a sector-native GAS space computes sigma in fp64 or fp32")
            e, c, conv, stats = solve(self, gas.cas, h2e, pre_diag, nroots, ci0, precision,
                                      tol=tol, tol_residual=tol_residual,
                                      max_cycle=max_cycle, max_space=max_space,
                                      lindep=lindep, level_shift=level_shift,
                                      verbose=verbose, basis=basis,
                                      follow_state=penalty is None)
            self.stats = stats
        e = numpy.asarray(e, dtype=numpy.float64).reshape(-1) + float(ecore)
        if penalty is not None:
            pen = numpy.array([penalty.value(x.reshape(-1),
                                             spin.contract_dev(x.reshape(-1),
                                                               torch.empty_like(x.reshape(-1))))
                               for x in c[:nroots]])
            phys = e[:nroots] - pen
            self.e_spin_penalty = float(pen[0]) if nroots == 1 else pen
            self.e_physical = float(phys[0]) if nroots == 1 else phys
        cis = [x.reshape(-1).cpu().numpy() for x in c]
        if nroots == 1:
            self.converged = bool(numpy.asarray(conv).reshape(-1)[0])
            self.eci = float(e[0])
            self.ci = cis[0]
        else:
            self.converged = numpy.asarray(conv, dtype=bool)[:nroots]
            self.eci = e[:nroots]
            self.ci = cis[:nroots]
        return self.eci, self.ci

    def _penalty_guesses(self, gas, spin, penalty, hdiag, ci0, nroots, lindep):
        """Initial vectors of a penalised solve (fci_gas.py:1269-1338): the
        default or given guesses projected on the target spin and
        orthonormalised (nroots of them), then the unprojected guesses of
        the first roots as global probes."""
        import torch
        target, eig = penalty.target, penalty.spectrum
        count = min(gas.ndet, max(4 * nroots, nroots + 4))
        default = self._init_guess_dev(hdiag, count)
        if ci0 is None:
            guess = default
        else:
            guess = [torch.from_numpy(v).cuda() for v in self._guess_list(ci0, nroots, gas.ndet)]
            if len(guess) < count:
                guess = guess + default[len(guess):count]
        projected = []
        for cand in guess:
            v = spin.project(cand, target, eig)
            for prev in projected:
                v = v - torch.dot(prev, v) * prev
            nrm = float(torch.linalg.norm(v))
            if nrm > max(100.0 * lindep, 1e-12):
                projected.append(v / nrm)
            if len(projected) == nroots:
                break
        if len(projected) < nroots:
            raise RuntimeError("could not construct enough independent target-spin "
                               "initial guesses")
        probes = [g / torch.linalg.norm(g) for g in guess[:max(1, nroots)]
                  if float(torch.linalg.norm(g)) > 0.0]
        return projected + probes

    @staticmethod
    def _guess_list(ci0, nroots, ndet):
        """Normalised initial vectors from a list, an (nroots, ndet) or
        (ndet, nroots) array, or one vector (fci_gas.py:1275-1305)."""
        if isinstance(ci0, (list, tuple)):
            vecs = [numpy.asarray(x, dtype=numpy.float64).reshape(-1) for x in ci0]
        else:
            arr = numpy.asarray(ci0, dtype=numpy.float64)
            if arr.ndim == 2 and arr.shape[0] == nroots and arr.shape[1] == ndet:
                vecs = list(arr)
            elif arr.ndim == 2 and arr.shape[1] == nroots and arr.shape[0] == ndet:
                vecs = list(arr.T)
            else:
                vecs = [arr.reshape(-1)]
        out = []
        for v in vecs:
            if v.size != ndet:
                raise ValueError("initial CI vector size does not match GAS determinant count")
            nrm = numpy.linalg.norm(v)
            if nrm == 0:
                raise ValueError("initial CI vector must be nonzero")
            out.append(v / nrm)
        return out

    # -- densities -----------------------------------------------------------
    def _rdm_plan(self, norb, nelec, plan):
        if plan is None:
            return GasRDMPlan(self, norb, nelec, gas=self._space(norb, nelec))
        self._check_plan(plan, norb, nelec, "RDM", GasRDMPlan)
        return plan

    def make_rdm1s(self, ci, norb, nelec, link_index=None, *, plan=None):
        return self._rdm_plan(norb, nelec, plan).make_rdm1s(ci, ci)

    def make_rdm1(self, ci, norb, nelec, link_index=None, *, plan=None):
        return self._rdm_plan(norb, nelec, plan).make_rdm1(ci, ci)

    def make_rdm12s(self, ci, norb, nelec, link_index=None, reorder=True, *, plan=None):
        if not reorder:
            raise NotImplementedError("reorder=False is not supported")
        return self._rdm_plan(norb, nelec, plan).make_rdm12s(ci, ci)

    def make_rdm12(self, ci, norb, nelec, link_index=None, reorder=True, *, plan=None):
        if not reorder:
            raise NotImplementedError("reorder=False is not supported")
        return self._rdm_plan(norb, nelec, plan).make_rdm12(ci, ci)

    def make_rdm2(self, ci, norb, nelec, link_index=None, reorder=True, *, plan=None):
        return self.make_rdm12(ci, norb, nelec, link_index, reorder, plan=plan)[1]

    def make_rdm123(self, *args, **kwargs):
        raise NotImplementedError("third- and fourth-order density matrices are not "
                                  "implemented for the GAS solver")

    make_rdm123s = make_rdm1234 = make_rdm1234s = make_rdm123

    def trans_rdm1s(self, cibra, ciket, norb, nelec, link_index=None, *, plan=None):
        return self._rdm_plan(norb, nelec, plan).make_rdm1s(cibra, ciket)

    def trans_rdm1(self, cibra, ciket, norb, nelec, link_index=None, *, plan=None):
        return self._rdm_plan(norb, nelec, plan).make_rdm1(cibra, ciket)

    def trans_rdm12s(self, cibra, ciket, norb, nelec, link_index=None, reorder=True, *,
                     plan=None):
        if not reorder:
            raise NotImplementedError("reorder=False is not supported")
        p = self._rdm_plan(norb, nelec, plan)
        dm1s, (aa, ab, bb) = p.make_rdm12s(cibra, ciket)
        _, (_, ba_ji, _) = p.make_rdm12s(ciket, cibra)
        return dm1s, (aa, ab, ba_ji.transpose(3, 2, 1, 0), bb)

    def trans_rdm12(self, cibra, ciket, norb, nelec, link_index=None, reorder=True, *,
                    plan=None):
        if not reorder:
            raise NotImplementedError("reorder=False is not supported")
        return self._rdm_plan(norb, nelec, plan).make_rdm12(cibra, ciket)

    def spin_square(self, ci, norb, nelec, *args, **kwargs):
        nelec = _ds._unpack_nelec(nelec, self.spin)
        dm1s, dm2s = self.make_rdm12s(ci, norb, nelec)
        return _ds.spin_square_from_rdm12s(dm1s, dm2s, nelec)

    def transform_ci_within_gas(self, fcivec, norb, nelec, u):
        """CI coefficients for mo_new = mo_old @ u, u orthogonal and block
        diagonal over the GAS subspaces (fci_gas.py:1003-1091): each block
        u_g = G_1^T ... G_m^T D (adjacent-orbital Givens rotations, D = +-1)
        acts on the string axes of subspace g of every block as the
        rotations in order, then the signs of D; on the device
        (fcis_rotate_fibers), one launch per (subspace, electron count,
        spin) axis.  numpy in, numpy out; the input is not modified."""
        import torch
        orbs, ne, _ = self._space_spec(norb, nelec)
        norb = int(norb)
        u = numpy.asarray(u, dtype=numpy.float64)
        if u.shape != (norb, norb) or not numpy.all(numpy.isfinite(u)) or not numpy.allclose(
                u.T @ u, numpy.eye(norb), atol=1e-10, rtol=0):
            raise ValueError("GAS orbital rotation must be finite and orthogonal")
        bounds = numpy.cumsum((0,) + tuple(orbs))
        inside = numpy.zeros((norb, norb), dtype=bool)
        for g in range(len(orbs)):
            inside[bounds[g]:bounds[g + 1], bounds[g]:bounds[g + 1]] = True
        if numpy.any(numpy.abs(u[~inside]) > 1e-12):
            raise ValueError("orbital rotation must not mix different GAS subspaces")
        factors = [_givens_factors(u[bounds[g]:bounds[g + 1], bounds[g]:bounds[g + 1]])
                   for g in range(len(orbs))]
        gas = self._space(norb, nelec)
        c, host, shape = _vector_in(fcivec, gas.ndet)
        x = c.clone()
        occ, _ = gas.sector_occupations()
        G = len(orbs)
        axes = {}
        for blk in gas.block_descriptors():
            occs = [int(v) for v in occ[blk["sa"]]] + [int(v) for v in occ[blk["sb"]]]
            dims = [math.comb(orbs[a % G], occs[a]) for a in range(2 * G)]
            for a in range(2 * G):
                st = int(numpy.prod(dims[a + 1:], dtype=numpy.int64))
                outer = int(numpy.prod(dims[:a], dtype=numpy.int64))
                base = blk["offset"] + (numpy.arange(outer, dtype=numpy.int64)[:, None] * (dims[a] * st)
                                        + numpy.arange(st, dtype=numpy.int64)[None, :]).reshape(-1)
                axes.setdefault((a % G, occs[a], a // G), []).append((base, st))
        lib = _lib.load()
        for (g, k, _spin), fibers in axes.items():
            rots, d = factors[g]
            strings = numpy.array(sorted(sum(1 << p for p in o) for o in
                                         itertools.combinations(range(orbs[g]), k)), dtype=numpy.int64)
            sign = numpy.ones(strings.size)
            for p in numpy.flatnonzero(d < 0):
                sign *= 1 - 2 * ((strings >> p) & 1)
            if not rots and numpy.all(sign > 0):
                continue
            roff, lefts, rights, cs = [0], [], [], []
            for p, q, cc, ss in rots:
                left = numpy.flatnonzero(((strings >> p) & 1) & (1 - ((strings >> q) & 1)))
                lefts.append(left)
                rights.append(numpy.searchsorted(strings, strings[left] ^ ((1 << p) | (1 << q))))
                roff.append(roff[-1] + left.size)
                cs += [cc, ss]
            dev = lambda a, t: torch.as_tensor(numpy.ascontiguousarray(a, dtype=t)).cuda()  # noqa: E731
            base = dev(numpy.concatenate([f[0] for f in fibers]), numpy.int64)
            stride = dev(numpy.concatenate([numpy.full(f[0].size, f[1], dtype=numpy.int64)
                                            for f in fibers]), numpy.int64)
            # one trailing pad keeps the pair arrays non-empty (no pairs when the
            # subspace is empty or full); roff bounds every rotation's range
            cat = (lambda v: numpy.concatenate(v + [numpy.zeros(1, dtype=numpy.int64)]))
            t_roff = dev(roff, numpy.int32)
            t_left, t_right = dev(cat(lefts), numpy.int32), dev(cat(rights), numpy.int32)
            t_cs, t_sign = dev(cs if cs else [0.0], numpy.float64), dev(sign, numpy.float64)
            _lib.check(lib.fcis_rotate_fibers(_lib.vptr(x), int(base.numel()), _lib.vptr(base),
                                              _lib.vptr(stride), int(strings.size), len(rots),
                                              _lib.vptr(t_roff), _lib.vptr(t_left),
                                              _lib.vptr(t_right), _lib.vptr(t_cs),
                                              _lib.vptr(t_sign), _lib.stream_ptr()),
                       "fcis_rotate_fibers")
            torch.cuda.current_stream().synchronize()
        return x.cpu().numpy().reshape(shape) if host else x.reshape(shape)

    def transform_ci_for_orbital_rotation(self, fcivec, norb, nelec, u):
        raise NotImplementedError("arbitrary active-orbital rotations do not preserve a GAS space")

    def dump_flags(self, verbose=None):
        """FCI settings and the spin-penalty ones (fci_gas.py:805-814)."""
        log = print if (verbose if verbose is not None else self.verbose) else (lambda *a: None)
        log("******** %s ********" % self.__class__)
        log("gas_orbs = %s, gas_restr_type = %s" % (self.gas_orbs, self.gas_restr_type))
        log("max. cycles = %d, conv_tol = %g, max_space = %d, precision = %s" % (
            self.max_cycle, self.conv_tol, self.max_space, self.precision))
        if getattr(self, "ss_penalty", None) is not None:
            target = getattr(self, "ss_value", None)
            log("spin penalty shift = %g" % self.ss_penalty)
            log("target S^2 = %s" % ("minimum" if target is None else target))
        return self

    def to_gpu(self, *args, **kwargs):
        """Already on the GPU."""
        return self


spin_square_from_rdm12s = _ds.spin_square_from_rdm12s

"""Spin-adapted (CSF-basis) CAS FCI on the GPU.

Mirrors pyscf-forge's CSF-FCI solver (pyscf/csf_fci/csf.py) and its
transformer (pyscf/csf_fci/csfstring.py:33-120):

* ``CSFTransformer(norb, neleca, nelecb, smult)`` with ``vec_det2csf`` /
  ``vec_csf2det`` (csfstring.py:45-60; same argument meaning, normalisation
  and return_norm behaviour), ``ncsf``, ``ndet``.  The transform runs in
  HIP kernels (fcis_csf_det2csf / fcis_csf_csf2det); host numpy vectors are
  copied to HBM and back, device torch tensors stay on the device.
* ``make_hdiag_csf`` (csf.py:70-144) on the device.
* ``FCISolver`` (csf.py:499-560): attributes ``smult`` and ``transformer``;
  ``kernel`` follows csf.kernel (csf.py:368-497): Davidson in the CSF basis
  with hop = det2csf . sigma . csf2det (csf.py:447-450), preconditioned by
  the CSF diagonal, started from the lowest-diagonal CSFs
  (get_init_guess csf.py:51-64), exact diagonalisation when the CSF space
  fits the pspace (csf.py:416-429); ``ci0`` and the returned ``ci`` are
  determinant-basis (na, nb) arrays.
* Point-group symmetry (csf_symm.py, csfstring.py:74-103): with ``orbsym``
  (D2h-subgroup irrep ids, products by XOR) and ``wfnsym``, CSF vectors are
  packed to the configurations of irrep ``wfnsym`` (fcis_csf_pack /
  _unpack on the device) and the solver works in that subspace.
"""

import itertools
import math

import numpy

from . import _lib
from . import direct_spin1 as _ds


class ImpossibleSpinError(RuntimeError):
    """No CSF of the requested multiplicity (csfstring.py:27-31)."""


class CSFTransformer:
    """Determinant <-> CSF transform for (norb, neleca, nelecb, 2S+1)."""

    def __init__(self, norb, neleca, nelecb, smult, orbsym=None, wfnsym=None, max_memory=None):
        self.max_memory = max_memory
        self._h = None
        self._orbsym = None
        self._wfnsym = None
        self.nsym = None
        self._update_spin_cache(norb, neleca, nelecb, smult)
        self._update_symm_cache(orbsym, wfnsym)

    # -- point group (csfstring.py:74-103) -------------------------------
    @property
    def orbsym(self):
        return self._orbsym

    @property
    def wfnsym(self):
        return self._wfnsym

    @wfnsym.setter
    def wfnsym(self, value):
        self._update_symm_cache(self._orbsym, value)

    def _update_symm_cache(self, orbsym, wfnsym=None):
        """Pack to the CSFs whose configuration irrep is wfnsym (both
        given), or clear the packing."""
        import ctypes
        orbsym = None if orbsym is None else tuple(int(x) for x in orbsym)
        wfnsym = None if wfnsym is None else int(wfnsym)
        if orbsym is not None and len(orbsym) != self.norb:
            raise ValueError("orbsym needs one irrep per orbital")
        self._orbsym, self._wfnsym = orbsym, wfnsym
        lib = _lib.load()
        n = ctypes.c_int64()
        if orbsym is None or wfnsym is None:
            _lib.check(lib.fcis_csf_symmetry(self._h, None, 0, ctypes.byref(n), _lib.stream_ptr()),
                       "fcis_csf_symmetry")
            self.nsym = None
            return
        arr = numpy.ascontiguousarray(orbsym, dtype=numpy.int32)
        _lib.check(lib.fcis_csf_symmetry(self._h, arr.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                         wfnsym, ctypes.byref(n), _lib.stream_ptr()),
                   "fcis_csf_symmetry")
        self.nsym = int(n.value)

    def csf_irrep(self, csf_index, orbsym=None):
        """Irrep of the configuration of (full-space) CSF ``csf_index``."""
        import ctypes
        arr = numpy.ascontiguousarray(orbsym if orbsym is not None else self._orbsym,
                                      dtype=numpy.int32)
        out = ctypes.c_int()
        _lib.check(_lib.load().fcis_csf_irrep(self._h, arr.ctypes.data_as(
            ctypes.POINTER(ctypes.c_int)), int(csf_index), ctypes.byref(out)), "fcis_csf_irrep")
        return int(out.value)

    @property
    def symmetric(self):
        return self.nsym is not None

    @property
    def ncsf_sym(self):
        """Length of the (packed) CSF vectors of this transformer."""
        return self.nsym if self.symmetric else self.ncsf_full

    @property
    def ncsf(self):
        """CSF count, of irrep wfnsym under symmetry (csfstring.py:210-213);
        ncsf_full counts every irrep."""
        return self.ncsf_sym

    def pack_dev(self, full, out, nvec=1):
        _lib.check(_lib.load().fcis_csf_pack(self._h, _lib.vptr(full), _lib.vptr(out), int(nvec),
                                             _lib.stream_ptr()), "fcis_csf_pack")

    def unpack_dev(self, packed, out, nvec=1):
        _lib.check(_lib.load().fcis_csf_unpack(self._h, _lib.vptr(packed), _lib.vptr(out),
                                               int(nvec), _lib.stream_ptr()), "fcis_csf_unpack")

    # -- cache ---------------------------------------------------------
    def _update_spin_cache(self, norb, neleca, nelecb, smult):
        key = (int(norb), int(neleca), int(nelecb), int(smult))
        if self._h is not None and key == self._key:
            return
        self.close()
        import ctypes
        self._space = _ds.get_space(norb, (neleca, nelecb))
        lib = _lib.load()
        h = ctypes.c_void_p()
        st = lib.fcis_csf_create(ctypes.byref(h), self._space.handle, key[3], _lib.stream_ptr())
        if st != _lib.SUCCESS:
            msg = lib.fcis_last_error().decode()
            if st == _lib.ERR_INVALID:
                raise ImpossibleSpinError(msg)
            _lib.check(st, "fcis_csf_create")
        self._h = h
        self._key = key
        self.nsym = None
        ncsf = ctypes.c_int64()
        ndet = ctypes.c_int64()
        _lib.check(lib.fcis_csf_info(h, ctypes.byref(ncsf), ctypes.byref(ndet)), "fcis_csf_info")
        self.ncsf_full = int(ncsf.value)
        self.ndet = int(ndet.value)

    @property
    def norb(self):
        return self._key[0]

    @property
    def neleca(self):
        return self._key[1]

    @property
    def nelecb(self):
        return self._key[2]

    @property
    def smult(self):
        return self._key[3]

    @property
    def ndeta(self):
        return self._space.shape[0]

    @property
    def ndetb(self):
        return self._space.shape[1]

    @property
    def n(self):
        """Davidson basis size: packed CSF count under symmetry."""
        return self.ncsf_sym

    def close(self):
        if getattr(self, "_h", None) is not None:
            _lib.load().fcis_csf_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- device transforms (no normalisation) --------------------------
    def det2csf_dev(self, det, out, nvec=1):
        """out[nvec, ncsf] <- det[nvec, ndet] (device float64, contiguous)."""
        _lib.check(_lib.load().fcis_csf_det2csf(self._h, _lib.vptr(det), _lib.vptr(out),
                                                int(nvec), _lib.stream_ptr()),
                   "fcis_csf_det2csf")

    def csf2det_dev(self, csf, out, nvec=1):
        """out[nvec, ndet] <- csf[nvec, ncsf] (device float64, contiguous)."""
        _lib.check(_lib.load().fcis_csf_csf2det(self._h, _lib.vptr(csf), _lib.vptr(out),
                                                int(nvec), _lib.stream_ptr()),
                   "fcis_csf_csf2det")

    def hop(self, plan):
        """x (CSF, packed under symmetry) -> H x through the
        determinant-space sigma plan."""
        import torch
        det = torch.empty(self.ndet, dtype=torch.float64, device="cuda")
        sig = torch.empty_like(det)
        if not self.symmetric:
            def apply(x, out):
                self.csf2det_dev(x, det)
                plan(det, sig)
                self.det2csf_dev(sig, out)
            return apply
        full = torch.empty(self.ncsf_full, dtype=torch.float64, device="cuda")

        def apply_sym(x, out):
            self.unpack_dev(x, full)
            self.csf2det_dev(full, det)
            plan(det, sig)
            self.det2csf_dev(sig, full)
            self.pack_dev(full, out)
        return apply_sym

    # -- pyscf-facing transforms ----------------------------------------
    def _apply(self, civec, nin, nout, fn, order, normalize, return_norm):
        import torch
        on_dev = isinstance(civec, torch.Tensor)
        x = civec if on_dev else torch.as_tensor(numpy.asarray(civec, dtype=numpy.float64))
        x = x.to(device="cuda", dtype=torch.float64)
        vec_on_cols = order.upper() == "F"
        if x.numel() == nin:
            x = x.reshape(1, nin)
            single = True
        else:
            x = x.reshape(nin, -1).T if vec_on_cols else x.reshape(-1, nin)
            single = False
        x = x.contiguous()
        nvec = x.shape[0]
        out = torch.empty((nvec, nout), dtype=torch.float64, device="cuda")
        if nout:
            fn(x, out, nvec)
        norm = torch.linalg.norm(out, dim=1)
        if normalize:
            out = out / norm[:, None]
        if vec_on_cols and not single:
            out = out.T.contiguous()
        if single:
            out = out.reshape(nout)
            norm = norm[0]
        if not on_dev:
            out = out.cpu().numpy()
            norm = norm.cpu().numpy() if not single else float(norm)
        return (out, norm) if return_norm else out

    def _det2csf_packed(self, det, out, nvec):
        import torch
        full = torch.empty((nvec, self.ncsf_full), dtype=torch.float64, device="cuda")
        self.det2csf_dev(det, full, nvec)
        self.pack_dev(full, out, nvec)

    def _csf2det_packed(self, csf, out, nvec):
        import torch
        full = torch.empty((nvec, self.ncsf_full), dtype=torch.float64, device="cuda")
        self.unpack_dev(csf, full, nvec)
        self.csf2det_dev(full, out, nvec)

    def vec_det2csf(self, civec, order="C", normalize=True, return_norm=False):
        """Determinant CI vector(s) -> CSF vector(s), packed to wfnsym under
        symmetry (csfstring.py:45-52)."""
        fn = self._det2csf_packed if self.symmetric else self.det2csf_dev
        return self._apply(civec, self.ndet, self.ncsf_sym, fn, order, normalize, return_norm)

    def vec_csf2det(self, civec, order="C", normalize=True, return_norm=False):
        """CSF vector(s) (packed under symmetry) -> flat determinant
        vector(s) (csfstring.py:54-60)."""
        fn = self._csf2det_packed if self.symmetric else self.csf2det_dev
        return self._apply(civec, self.ncsf_sym, self.ndet, fn, order, normalize, return_norm)

    def _pack(self, csfvec, order, nin, nout, fn):
        import torch
        on_dev = isinstance(csfvec, torch.Tensor)
        x = csfvec if on_dev else torch.as_tensor(numpy.asarray(csfvec, dtype=numpy.float64))
        x = x.to(device="cuda", dtype=torch.float64)
        cols = order.upper() == "F"
        single = x.numel() == nin
        x = x.reshape(1, nin) if single else (x.reshape(nin, -1).T if cols else x.reshape(-1, nin))
        x = x.contiguous()
        out = torch.empty((x.shape[0], nout), dtype=torch.float64, device="cuda")
        fn(x, out, x.shape[0])
        out = out.reshape(nout) if single else (out.T.contiguous() if cols else out)
        return out if on_dev else out.cpu().numpy()

    def pack_csf(self, csfvec, order="C"):
        """Full CSF vector(s) -> the wfnsym subspace (csfstring.py:74-80)."""
        if not self.symmetric:
            return csfvec
        return self._pack(csfvec, order, self.ncsf_full, self.nsym, self.pack_dev)

    def unpack_csf(self, csfvec, order="C"):
        """wfnsym-packed CSF vector(s) -> full, zero elsewhere (:82-88)."""
        if not self.symmetric:
            return csfvec
        return self._pack(csfvec, order, self.nsym, self.ncsf_full, self.unpack_dev)

    # -- determinant packing, projections, printing (csfstring.py:42-220) --
    def _det_sym_index(self):
        """Device addresses of the determinants whose configuration has
        irrep wfnsym (XOR of orbsym over the singly occupied orbitals, the
        same as over every occupied spin orbital), ascending."""
        import torch
        key = (self._orbsym, self._wfnsym, self._key)
        if getattr(self, "_det_idx_key", None) != key:
            def irreps(k):
                occ = numpy.array(list(itertools.combinations(range(self.norb), k)), dtype=numpy.int64)
                occ = occ.reshape(-1, k)
                val = (numpy.int64(1) << occ).sum(axis=1) if k else numpy.zeros(1, numpy.int64)
                occ = occ[numpy.argsort(val, kind="stable")]     # colex = string value order
                sym = numpy.asarray(self._orbsym, dtype=numpy.int64)
                out = numpy.zeros(occ.shape[0], dtype=numpy.int64)
                for col in range(k):
                    out ^= sym[occ[:, col]]
                return out
            ia, ib = irreps(self.neleca), irreps(self.nelecb)
            mask = (ia[:, None] ^ ib[None, :]).reshape(-1) == self._wfnsym
            self._det_idx = torch.from_numpy(numpy.nonzero(mask)[0]).cuda()
            self._det_idx_key = key
        return self._det_idx

    def _rows_in(self, vec, order, nin):
        import torch
        on_dev = isinstance(vec, torch.Tensor)
        x = vec if on_dev else torch.as_tensor(numpy.asarray(vec, dtype=numpy.float64))
        x = x.to(device="cuda", dtype=torch.float64)
        cols = order.upper() == "F"
        single = x.numel() == nin
        x = x.reshape(1, nin) if single else (x.reshape(nin, -1).T if cols else x.reshape(-1, nin))
        return x, on_dev, cols, single

    @staticmethod
    def _rows_out(out, on_dev, cols, single):
        out = out.reshape(-1) if single else (out.T.contiguous() if cols else out)
        return out if on_dev else out.cpu().numpy()

    def pack_det(self, detvec, order="C"):
        """Determinant vector(s) -> the determinants of irrep wfnsym
        (csfstring.py:88-96); unchanged without symmetry."""
        if not self.symmetric:
            return detvec
        x, on_dev, cols, single = self._rows_in(detvec, order, self.ndet)
        return self._rows_out(x[:, self._det_sym_index()].contiguous(), on_dev, cols, single)

    def unpack_det(self, detvec, order="C"):
        """Packed determinant vector(s) -> full, zero elsewhere
        (csfstring.py:98-103)."""
        import torch
        if not self.symmetric:
            return detvec
        idx = self._det_sym_index()
        x, on_dev, cols, single = self._rows_in(detvec, order, idx.numel())
        out = torch.zeros((x.shape[0], self.ndet), dtype=torch.float64, device="cuda")
        out[:, idx] = x
        return self._rows_out(out, on_dev, cols, single)

    def project_civec(self, detarr, order="C", normalize=True, return_norm=False):
        """Projection of determinant vector(s) on this spin (and irrep):
        vec_csf2det(vec_det2csf(x)).  The reference leaves this
        unimplemented (csfstring.py:42-43); ``return_norm`` gives the norm
        of the projection before normalisation."""
        x, norm = self.vec_det2csf(detarr, order=order, normalize=False, return_norm=True)
        out = self.vec_csf2det(x, order=order, normalize=normalize)
        return (out, norm) if return_norm else out

    def mat_det2csf(self, mat):
        raise NotImplementedError

    def mat_csf2det(self, mat):
        raise NotImplementedError

    def full_index_dev(self, idx):
        """Full-space CSF indices (device int64) of packed ones (identity
        without symmetry)."""
        import torch
        idx = torch.as_tensor(idx).to(device="cuda", dtype=torch.int64).reshape(-1)
        if not self.symmetric:
            return idx
        key = (self._key, self._orbsym, self._wfnsym)
        if getattr(self, "_sym_full_key", None) != key:
            pos = torch.arange(1, self.nsym + 1, dtype=torch.float64, device="cuda")
            full = torch.empty(self.ncsf_full, dtype=torch.float64, device="cuda")
            self.unpack_dev(pos, full)
            self._sym_full = torch.nonzero(full).reshape(-1)
            self._sym_full_key = key
        return self._sym_full[idx]

    def _full_index(self, idx):
        idx = numpy.atleast_1d(numpy.asarray(idx, dtype=numpy.int64))
        return self.full_index_dev(idx).cpu().numpy()

    @property
    def max_support(self):
        """Largest number of determinants in one configuration."""
        import ctypes
        v = ctypes.c_int()
        _lib.check(_lib.load().fcis_csf_max_support(self._h, ctypes.byref(v)),
                   "fcis_csf_max_support")
        return v.value

    def supports(self, idx):
        """Determinant supports of CSFs ``idx`` (packed under symmetry):
        device (addr [m, maxdet], coef [m, maxdet], count [m]), row i the
        nonzero entries of CSF i's determinant expansion."""
        import torch
        full = self.full_index_dev(idx)
        m, mx = int(full.numel()), self.max_support
        A = torch.empty((m, mx), dtype=torch.int64, device="cuda")
        C = torch.empty((m, mx), dtype=torch.float64, device="cuda")
        n = torch.empty(m, dtype=torch.int32, device="cuda")
        _lib.check(_lib.load().fcis_csf_support(self._h, m, _lib.vptr(full), mx, _lib.vptr(A),
                                                _lib.vptr(C), _lib.vptr(n), _lib.stream_ptr()),
                   "fcis_csf_support")
        return A, C, n

    def printable_csfstring(self, idx):
        """'2'/'0'/'u'/'d' string (highest orbital first) of CSF(s) idx,
        packed indices under symmetry (csfstring.py:147-155)."""
        isarr = hasattr(idx, "__len__")
        ret = [_printable(self._h, self.norb, int(i)) for i in self._full_index(idx)]
        return ret if isarr else ret[0]

    def printable_largest_csf(self, csfvec, npr, order="C", isdet=False, normalize=True):
        """(strings, coefficients) of the npr largest |coefficients| of each
        vector (csfstring.py:122-145)."""
        if isdet:
            csfvec = self.vec_det2csf(csfvec, order=order, normalize=normalize)
        csfvec = self.unpack_csf(csfvec, order=order)
        arr = numpy.asarray(csfvec.cpu().numpy() if hasattr(csfvec, "cpu") else csfvec)
        if arr.ndim == 1:
            vecs = [arr]
        elif order.upper() == "C":
            vecs = [arr[i, :] for i in range(arr.shape[0])]
        elif order.upper() == "F":
            vecs = [arr[:, i] for i in range(arr.shape[1])]
        else:
            raise RuntimeError('order must be "C" or "F"')
        npr = min(vecs[0].size, npr)
        idx = [numpy.argsort(-numpy.abs(v))[:npr] for v in vecs]
        coef = [v[i] for v, i in zip(vecs, idx)]
        strs = [[_printable(self._h, self.norb, int(k)) for k in i] for i in idx]
        axis = int(order.upper() == "F")
        return numpy.stack(strs, axis=axis), numpy.stack(coef, axis=axis)

    def print_config(self, printer=print):
        printer("***** CSFTransformer configuration *****")
        printer("norb = {}".format(self.norb))
        printer("neleca, nelecb = {}, {}".format(self.neleca, self.nelecb))
        printer("smult = {}".format(self.smult))
        printer("orbsym = {}".format(self.orbsym))
        printer("wfnsym = {}".format(self.wfnsym))
        printer("ndeta, ndetb = {}, {}".format(self.ndeta, self.ndetb))
        printer("ncsf = {}".format(self.ncsf))


def _printable(handle, norb, csf):
    """One CSF as a string over the orbitals, highest first: '2' doubly
    occupied, '0' empty, 'u' / 'd' singly occupied coupled up / down
    (_printable_csfstring, csfstring.py:1187-1207)."""
    import ctypes
    d, so, cp = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
    _lib.check(_lib.load().fcis_csf_describe(handle, int(csf), ctypes.byref(d), ctypes.byref(so),
                                             ctypes.byref(cp)), "fcis_csf_describe")
    d, so, cp = d.value, so.value, cp.value
    chars = []
    k = 0
    for p in range(norb):
        if (d >> p) & 1:
            chars.append("2")
        elif (so >> p) & 1:
            chars.append("u" if (cp >> k) & 1 else "d")
            k += 1
        else:
            chars.append("0")
    return "".join(reversed(chars))


def _with_transformer(norb, neleca, nelecb, smult, fn):
    t = CSFTransformer(norb, neleca, nelecb, smult)
    try:
        return fn(t)
    finally:
        t.close()


def transform_civec_det2csf(detarr, norb, neleca, nelecb, smult, csd_mask=None,
                            vec_on_cols=False, do_normalize=True, max_memory=None):
    """(csfarr, norm) (csfstring.py:341-414), on the device."""
    order = "F" if vec_on_cols else "C"
    return _with_transformer(norb, neleca, nelecb, smult, lambda t: t.vec_det2csf(
        detarr, order=order, normalize=do_normalize, return_norm=True))


def transform_civec_csf2det(csfarr, norb, neleca, nelecb, smult, csd_mask=None,
                            vec_on_cols=False, do_normalize=True, max_memory=None):
    """(detarr, norm) (csfstring.py:416-487), on the device."""
    order = "F" if vec_on_cols else "C"
    return _with_transformer(norb, neleca, nelecb, smult, lambda t: t.vec_csf2det(
        csfarr, order=order, normalize=do_normalize, return_norm=True))


def check_spinstate_norm(detarr, norb, neleca, nelecb, smult, csd_mask=None):
    """Norm of the projection of detarr on 2S+1 = smult (csfstring.py:297-299)."""
    return transform_civec_det2csf(detarr, norb, neleca, nelecb, smult)[1]


def project_civec_csf(detarr, norb, neleca, nelecb, smult, csd_mask=None, max_memory=None):
    """(normalised projection on 2S+1 = smult in the determinant basis,
    its norm) (csfstring.py:302-339); the input shape is kept."""
    shape = numpy.shape(detarr)

    def fn(t):
        out, norm = t.project_civec(numpy.asarray(detarr).reshape(-1), normalize=True,
                                    return_norm=True)
        return out.reshape(shape), norm
    return _with_transformer(norb, neleca, nelecb, smult, fn)


def printable_csfstring(norb, neleca, nelecb, smult, addrs):
    """Strings of full-space CSF addresses (csfstring.py:1183-1185)."""
    t = CSFTransformer(norb, neleca, nelecb, smult)
    try:
        return [_printable(t._h, norb, int(a)) for a in numpy.atleast_1d(addrs)]
    finally:
        t.close()


def count_csfs(nspin, smult):
    """CSFs of nspin singly occupied orbitals coupled to 2S+1 = smult:
    C(nspin, nspin/2 - S) - C(nspin, nspin/2 - S - 1) (branching rule)."""
    twos = int(smult) - 1
    if twos < 0 or twos > nspin or (nspin - twos) % 2:
        return 0
    k = (nspin - twos) // 2
    return math.comb(nspin, k) - (math.comb(nspin, k - 1) if k > 0 else 0)


def count_all_csfs(norb, neleca, nelecb, smult):
    """Number of CSFs of multiplicity smult (csfstring.count_all_csfs)."""
    if abs(neleca - nelecb) > smult - 1:
        return 0
    total = 0
    for npair in range(max(0, neleca + nelecb - norb), min(neleca, nelecb) + 1):
        nspin = neleca + nelecb - 2 * npair
        total += (math.comb(norb, npair) * math.comb(norb - npair, nspin)
                  * count_csfs(nspin, smult))
    return total


# ---- spin-dependent one-electron Hamiltonians (csf.py:28-49, 66-68, 518-529)
def unpack_h1e_ab(h1e):
    """(h_alpha, h_beta) of a one- or two-component h1e (csf.py:28-32)."""
    h = numpy.asarray(h1e)
    if h.ndim == 3 and h.shape[0] == 2:
        return h1e[0], h1e[1]
    return h1e, h1e


def unpack_1RDM_ab(dm1):
    """(dm_alpha, dm_beta) of a spin-summed or two-component 1-RDM (csf.py:34-38)."""
    dm1 = numpy.asarray(dm1)
    if dm1.ndim == 3 and dm1.shape[0] == 2:
        return dm1[0], dm1[1]
    return dm1 / 2, dm1 / 2


def unpack_h1e_cs(h1e):
    """(charge, spin) parts (h_a + h_b)/2, (h_a - h_b)/2 (csf.py:40-44)."""
    h1e_a, h1e_b = unpack_h1e_ab(h1e)
    h1e_a = numpy.asarray(h1e_a, dtype=numpy.float64)
    h1e_b = numpy.asarray(h1e_b, dtype=numpy.float64)
    return (h1e_a + h1e_b) / 2.0, (h1e_a - h1e_b) / 2.0


def unpack_1RDM_cs(dm):
    """(dm_a + dm_b, dm_a - dm_b) (csf.py:46-48)."""
    dma, dmb = unpack_1RDM_ab(dm)
    return dma + dmb, dma - dmb


class _SpinTaggedPair(numpy.ndarray):
    """Absorbed pair matrix carrying the spin part h1e_s of a two-component
    h1e (the tag_array of csf.py:518-523)."""

    def __array_finalize__(self, obj):
        self.h1e_s = getattr(obj, "h1e_s", None)


def _spin_part(h1e, norb):
    """h1e_s of h1e as a (norb, norb) array, or None when it vanishes."""
    h1e_c, h1e_s = unpack_h1e_cs(h1e)
    h1e_s = numpy.ascontiguousarray(h1e_s, dtype=numpy.float64).reshape(norb, norb)
    return h1e_c, (h1e_s if numpy.any(h1e_s) else None)


def _string_diag_sums(norb, nelec, d):
    """sum of d[p] over the occupied orbitals of each string, colex order."""
    occ = numpy.array(list(itertools.combinations(range(norb), nelec)), dtype=numpy.int64)
    occ = occ.reshape(-1, nelec)
    if nelec == 0:
        return numpy.zeros(1)
    val = (numpy.int64(1) << occ).sum(axis=1)
    occ = occ[numpy.argsort(val, kind="stable")]
    return d[occ].sum(axis=1)


def _hdiag_det_dev(h1e, eri, norb, nelec):
    """Determinant diagonal for a one- or two-component h1e
    (direct_uhf.make_hdiag as csf.make_hdiag_det wraps it, csf.py:66-68):
    the spin-free diagonal of (h_a + h_b)/2 plus sum_{p in a} s_pp -
    sum_{p in b} s_pp, s = (h_a - h_b)/2."""
    import torch
    norb = int(norb)
    na, nb = _ds._unpack_nelec(nelec)
    h1e_c, h1e_s = _spin_part(h1e, norb)
    hd = _ds._make_hdiag_dev(h1e_c, eri, norb, (na, nb))
    if h1e_s is not None:
        d = numpy.diag(h1e_s)
        corr = (_string_diag_sums(norb, na, d)[:, None] - _string_diag_sums(norb, nb, d)[None, :])
        hd += torch.from_numpy(corr.reshape(-1)).to(hd.device)
    return hd


def _h1e_array(h1e, norb):
    """h1e as a (norb, norb) or (2, norb, norb) float64 array."""
    h = numpy.asarray(h1e, dtype=numpy.float64)
    if h.ndim == 3 and h.shape[0] == 2:
        return numpy.ascontiguousarray(h).reshape(2, norb, norb)
    return numpy.ascontiguousarray(h).reshape(norb, norb)


def make_hdiag_det(fci, h1e, eri, norb, nelec):
    """Determinant diagonal, one- or two-component h1e (csf.py:66-68)."""
    return _hdiag_det_dev(h1e, eri, norb, nelec).cpu().numpy()


class _SpinSigma:
    """Sigma plan plus the spin part of a two-component h1e:
    out = H_c x + sum_pq s_pq (E^a_pq - E^b_pq) x."""

    def __init__(self, plan, sp, h1e_s):
        import torch
        self.plan, self.sp = plan, sp
        self.hs = torch.from_numpy(numpy.ascontiguousarray(h1e_s)).cuda()
        self.mhs = -self.hs

    def __call__(self, x, out):
        self.plan(x, out)
        _ds.contract_1e_ab_dev(self.sp, self.hs, self.mhs, x, out, accumulate=True)
        return out


class _SpinBasis:
    """A Davidson basis (CSF transformer) whose H application includes the
    spin part of h1e."""

    def __init__(self, basis, sp, h1e_s):
        self.basis, self.sp, self.h1e_s = basis, sp, h1e_s
        self.n = basis.n

    def hop(self, plan):
        return self.basis.hop(_SpinSigma(plan, self.sp, self.h1e_s))


def make_hdiag_csf(h1e, eri, norb, nelec, transformer, hdiag_det=None, max_memory=None):
    """CSF-basis diagonal of H (csf.py:70-144) as a device tensor; h1e may
    be two-component (h_alpha, h_beta)."""
    import torch
    norb = int(norb)
    nelec = _ds._unpack_nelec(nelec)
    eri = numpy.ascontiguousarray(_ds.pair_matrix(eri, norb))
    if hdiag_det is None:
        hdiag_det = _hdiag_det_dev(h1e, eri, norb, nelec)
    elif not isinstance(hdiag_det, torch.Tensor):
        hdiag_det = torch.as_tensor(numpy.asarray(hdiag_det, dtype=numpy.float64)).cuda()
    hdiag_det = hdiag_det.reshape(-1).contiguous()
    out = torch.empty(transformer.ncsf_full, dtype=torch.float64, device="cuda")
    _lib.check(_lib.load().fcis_csf_hdiag(transformer._h, _lib.dptr(eri), _lib.vptr(hdiag_det),
                                          _lib.vptr(out), _lib.stream_ptr()), "fcis_csf_hdiag")
    return out


class FCISolver(_ds.FCISolver):
    """CSF-basis CAS FCI solver (csf_fci.csf_solver for CAS, csf.py:499-560).

    ``smult`` = 2S+1 (default: the lowest allowed by nelec and ``spin``).
    ``orbsym`` / ``wfnsym`` (csf_symm.FCISolver, csf_symm.py:25-113):
    restrict the solve to CSFs of irrep ``wfnsym`` (default: the irrep of
    the lowest-diagonal CSF).  ``ci0``/``ci`` are determinant-basis arrays;
    Davidson runs on CSF vectors (packed to ``wfnsym``) kept in HBM."""

    def __init__(self, mol=None, smult=None, orbsym=None, wfnsym=None):
        super().__init__(mol)
        self.smult = smult
        self.orbsym = orbsym
        self.wfnsym = wfnsym
        self.transformer = None
        self.pspace_size = 200

    def check_transformer_cache(self, norb=None, nelec=None, smult=None):
        norb = self.norb if norb is None else norb
        nelec = self._nelec(self.nelec if nelec is None else nelec)
        if smult is not None:
            self.smult = smult
        if self.smult is None:
            self.smult = abs(nelec[0] - nelec[1]) + 1
        if self.transformer is None:
            self.transformer = CSFTransformer(norb, nelec[0], nelec[1], self.smult)
        else:
            self.transformer._update_spin_cache(norb, nelec[0], nelec[1], self.smult)
        self.transformer._update_symm_cache(self.orbsym, self.wfnsym)
        return self.transformer

    def _packed_hdiag(self, t, hdiag_full):
        import torch
        if not t.symmetric:
            return hdiag_full
        out = torch.empty(t.nsym, dtype=torch.float64, device="cuda")
        t.pack_dev(hdiag_full, out)
        return out

    def log_transformer_cache(self, tverbose=0, **kwargs):
        """Print the transformer settings (csf.py:542-552); tverbose > 0
        prints, 0 is silent."""
        if kwargs:
            self.__dict__.update(kwargs)
            self.check_transformer_cache()
        if self.transformer is None:
            return
        self.transformer.print_config(print if tverbose else (lambda *a, **k: None))

    def print_transformer_cache(self, **kwargs):
        return self.log_transformer_cache(10, **kwargs)

    def make_hdiag(self, h1e, eri, norb, nelec, *args, **kwargs):
        """Determinant diagonal, one- or two-component h1e (csf.py:502)."""
        return make_hdiag_det(self, h1e, eri, norb, nelec)

    def absorb_h1e(self, h1e, eri, norb, nelec, fac=1):
        """Pair matrix of the charge part (h_a + h_b)/2, tagged with the spin
        part (h_a - h_b)/2 when it is nonzero (csf.py:518-523)."""
        h1e_c, h1e_s = _spin_part(h1e, int(norb))
        h2 = _ds.absorb_h1e(h1e_c, eri, norb, nelec, fac)
        if h1e_s is not None:
            h2 = h2.view(_SpinTaggedPair)
            h2.h1e_s = h1e_s
        return h2

    def contract_2e(self, eri, fcivec, norb, nelec, link_index=None, **kwargs):
        """Sigma, plus sum_pq s_pq (E^a_pq - E^b_pq) c for a spin-tagged
        pair matrix (csf.py:525-529)."""
        import torch
        hc = _ds.contract_2e(eri, fcivec, norb, nelec, link_index,
                             precision=kwargs.pop("precision", "fp64"))
        h1e_s = getattr(eri, "h1e_s", None)
        if h1e_s is not None:
            sp = _ds.get_space(norb, self._nelec(nelec))
            on_dev = isinstance(fcivec, torch.Tensor)
            x = (fcivec if on_dev else torch.as_tensor(numpy.asarray(fcivec, dtype=numpy.float64)))
            x = x.to(device="cuda", dtype=torch.float64).contiguous().reshape(-1)
            out = (hc if isinstance(hc, torch.Tensor)
                   else torch.from_numpy(numpy.ascontiguousarray(hc))).to("cuda").reshape(-1)
            _ds.contract_1e_ab_dev(sp, h1e_s, -h1e_s, x, out, accumulate=True)
            hc = out.reshape(numpy.shape(hc)) if on_dev else out.cpu().numpy().reshape(numpy.shape(hc))
        return hc

    def make_hdiag_csf(self, h1e, eri, norb, nelec, hdiag_det=None, smult=None, max_memory=None):
        """CSF diagonal, packed to wfnsym under symmetry."""
        self.norb, self.nelec = int(norb), self._nelec(nelec)
        t = self.check_transformer_cache(smult=smult)
        return self._packed_hdiag(t, make_hdiag_csf(h1e, eri, norb, nelec, t,
                                                    hdiag_det=hdiag_det))

    @staticmethod
    def _init_guess_dev(hdiag, nroots):
        """Unit CSF vectors at the lowest diagonal elements, the first root
        with +-1e-5 at its ends (direct_spin1._get_init_guess applied to
        the CSF diagonal, as csf.get_init_guess does, csf.py:51-64)."""
        import torch
        n = hdiag.numel()
        k = min(nroots, n)
        addrs = _ds.lowest_addresses(hdiag, k)
        out = []
        for a in addrs.tolist():
            g = torch.zeros(n, dtype=torch.float64, device=hdiag.device)
            g[a] = 1.0
            out.append(g)
        if n > 1:
            out[0][0] += 1e-5
            out[0][-1] -= 1e-5
        return out

    def get_init_guess(self, norb, nelec, nroots, hdiag_csf):
        """Determinant-basis initial guesses (csf.get_init_guess)."""
        import torch
        hd = hdiag_csf if isinstance(hdiag_csf, torch.Tensor) else torch.as_tensor(
            numpy.asarray(hdiag_csf, dtype=numpy.float64)).cuda()
        self.norb, self.nelec = int(norb), self._nelec(nelec)
        t = self.check_transformer_cache()
        return [t.vec_csf2det(g).cpu().numpy() for g in self._init_guess_dev(hd, nroots)]

    def pspace(self, h1e, eri, norb, nelec, hdiag_det=None, hdiag_csf=None, npsp=200, **kwargs):
        """(csf_addr, h0): H in the npsp CSFs of lowest diagonal (csf.pspace,
        csf.py:246-366): the determinants of their configurations, the
        Slater-Condon block over those on the GPU, transformed to the CSFs."""
        import torch
        self.norb, self.nelec = int(norb), self._nelec(nelec)
        t = self.check_transformer_cache(smult=kwargs.get("smult"))
        sp = _ds.get_space(norb, self.nelec)
        h1e = _h1e_array(h1e, sp.norb)
        eri = _ds.pair_matrix(eri, sp.norb)
        if hdiag_csf is None:
            hdiag_csf = self._packed_hdiag(t, make_hdiag_csf(h1e, eri, norb, self.nelec, t,
                                                             hdiag_det=hdiag_det))
        elif not isinstance(hdiag_csf, torch.Tensor):
            hdiag_csf = torch.as_tensor(numpy.asarray(hdiag_csf, dtype=numpy.float64)).cuda()
        addr = self._pspace_addr(t, hdiag_csf, npsp)
        return addr.cpu().numpy(), self._h_block(sp, t, h1e, eri, addr)

    # determinants of the pspace configurations: at most this many (the
    # reference bounds the same block by max_memory, csf.py:287-300)
    PSPACE_MAX_DET = 20000

    def _pspace_addr(self, t, hdiag_csf, npsp):
        """Lowest-diagonal CSF addresses (ascending address order), trimmed
        from the top of the diagonal until their configurations span at most
        PSPACE_MAX_DET determinants."""
        import torch
        npsp = min(int(npsp), t.ncsf_sym)
        order = _ds.lowest_k(hdiag_csf, npsp)
        cnt = self._support_sizes(t, order)
        keep = int(numpy.searchsorted(numpy.cumsum(cnt), self.PSPACE_MAX_DET, side="right"))
        keep = max(1, keep)
        return torch.sort(order[:keep]).values

    @staticmethod
    def _support_sizes(t, addr):
        return t.supports(addr)[2].cpu().numpy().astype(numpy.int64)

    def _h_block(self, sp, t, h1e, eri, addr):
        """H over the CSFs ``addr``: U^T H_det U with H_det the Slater-Condon
        block over the union of their determinant supports (fcis_pspace)."""
        import torch
        m = addr.numel()
        A, C, n = t.supports(addr)
        valid = torch.arange(A.shape[1], device=A.device)[None, :] < n[:, None].to(torch.int64)
        dets = torch.unique(A[valid])
        if dets.numel() > max(self.PSPACE_MAX_DET, 1):
            raise ValueError("pspace of %d CSFs spans %d determinants (limit %d)"
                             % (m, dets.numel(), self.PSPACE_MAX_DET))
        U = torch.zeros((dets.numel(), m), dtype=torch.float64, device="cuda")
        cols = torch.arange(m, device=A.device)[:, None].expand_as(A)
        U[torch.searchsorted(dets, A[valid]), cols[valid]] = C[valid]
        h0 = _ds._pspace_dev(sp, h1e, eri, dets)
        H = (U.T @ (h0 @ U)).cpu().numpy()
        return 0.5 * (H + H.T)

    def guess_wfnsym(self, norb, nelec, fcivec=None, orbsym=None, wfnsym=None):
        """Target irrep id (direct_spin1_symm.guess_wfnsym, as csf_symm.py:66
        uses it): ``wfnsym`` if set, else the irrep of the largest
        determinant of ``fcivec``, else that of the aufbau determinant, the
        XOR of orbsym over its singly occupied orbitals.  None without
        orbsym."""
        orbsym = self.orbsym if orbsym is None else orbsym
        wfnsym = self.wfnsym if wfnsym is None else wfnsym
        if orbsym is None:
            return None
        if wfnsym is not None:
            if isinstance(wfnsym, str):
                raise TypeError("wfnsym must be an irrep id (int); irrep names need a "
                                "point-group table")
            return int(wfnsym)
        orbsym = [int(x) for x in orbsym]
        na, nb = self._nelec(nelec)
        if fcivec is None:
            w = 0
            for i in range(min(na, nb), max(na, nb)):
                w ^= orbsym[i]
            return w
        c = fcivec[0] if isinstance(fcivec, (list, tuple)) else fcivec
        sp = _ds.get_space(norb, (na, nb))
        c = numpy.asarray(c, dtype=numpy.float64).reshape(sp.shape)
        ia, ib = numpy.unravel_index(int(numpy.argmax(abs(c))), sp.shape)
        lib = _lib.load()
        occ = int(lib.fcis_addr2str(int(norb), na, int(ia)))
        occb = int(lib.fcis_addr2str(int(norb), nb, int(ib)))
        w = 0
        for i in range(int(norb)):
            if ((occ ^ occb) >> i) & 1:
                w ^= orbsym[i]
        return w

    def kernel(self, h1e, eri, norb, nelec, ci0=None, ecore=0, nroots=None, smult=None,
               **kwargs):
        """(e + ecore, ci) with ci in the determinant basis (csf.kernel;
        with ``orbsym``, csf_symm.FCISolver.kernel, csf_symm.py:33-77:
        orbsym / wfnsym kwargs apply to this call only)."""
        orbsym_back, wfnsym_back = self.orbsym, self.wfnsym
        if kwargs.get("orbsym") is not None:
            self.orbsym = kwargs["orbsym"]
        if kwargs.get("wfnsym") is not None:
            self.wfnsym = kwargs["wfnsym"]
        kwargs.pop("orbsym", None)
        kwargs.pop("wfnsym", None)
        try:
            self.wfnsym = self.guess_wfnsym(norb, nelec, ci0)
            return self._kernel(h1e, eri, norb, nelec, ci0, ecore, nroots, smult, **kwargs)
        finally:
            self.orbsym, self.wfnsym = orbsym_back, wfnsym_back

    def _kernel(self, h1e, eri, norb, nelec, ci0, ecore, nroots, smult, **kwargs):
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
        t = self.check_transformer_cache(smult=smult)
        sp = _ds.get_space(norb, nelec)
        nroots = min(nroots, t.ncsf_sym)
        if t.ncsf_sym == 0:
            raise ImpossibleSpinError("no CSF with 2S+1 = %d%s" % (
                self.smult, "" if not t.symmetric else " and irrep %d" % t.wfnsym))
        # a two-component h1e: the charge part is absorbed, the spin part
        # enters sigma as a one-electron term (csf.py:518-529)
        h1e = _h1e_array(h1e, sp.norb)
        eri = _ds.pair_matrix(eri, sp.norb)
        h1e_c, h1e_s = _spin_part(h1e, sp.norb)
        h2e = _ds.absorb_h1e(h1e_c, eri, norb, nelec, 0.5)
        hdiag_det = _hdiag_det_dev(h1e, eri, norb, nelec)
        hdiag = self._packed_hdiag(t, make_hdiag_csf(h1e, eri, norb, nelec, t,
                                                     hdiag_det=hdiag_det))

        # pspace of the lowest-diagonal CSFs (csf.py:414-420): the whole space
        # when it is small enough, else the block of the preconditioner
        npsp = max(pspace_size, nroots)
        if pspace_size >= t.ncsf_sym:
            paddr = torch.arange(t.ncsf_sym, device="cuda")
        elif pspace_size > 0:
            paddr = self._pspace_addr(t, hdiag, npsp)
        else:
            paddr = None
        pw = pv = None
        if paddr is not None:
            pw, pv = numpy.linalg.eigh(self._h_block(sp, t, h1e, eri, paddr))
        if ci0 is None and not davidson_only and pspace_size >= t.ncsf_sym:
            e = pw[:nroots]
            c = [torch.from_numpy(numpy.ascontiguousarray(pv[:, k])).cuda() for k in range(nroots)]
            conv = numpy.ones(nroots, dtype=bool)
        else:
            if ci0 is not None:
                vecs = list(ci0) if isinstance(ci0, (list, tuple)) else (
                    [ci0] if numpy.asarray(ci0).size == t.ndet else list(ci0))
                ci0 = [t.vec_det2csf(torch.as_tensor(numpy.asarray(v, dtype=numpy.float64))
                                     .cuda().reshape(-1)) for v in vecs]
            from .mixed import solve
            from .davidson import make_pspace_precond
            precond = None if pw is None else make_pspace_precond(hdiag, pw, pv, paddr,
                                                                   level_shift)
            e, c, conv, stats = solve(self, sp, h2e, hdiag, nroots, ci0, precision,
                                      tol=tol, tol_residual=tol_residual,
                                      max_cycle=max_cycle, max_space=max_space,
                                      lindep=lindep, level_shift=level_shift,
                                      verbose=verbose,
                                      basis=t if h1e_s is None else _SpinBasis(t, sp, h1e_s),
                                      precond=precond)
            self.stats = stats
        e = numpy.asarray(e, dtype=numpy.float64).reshape(-1) + float(ecore)
        cis = [t.vec_csf2det(x.reshape(-1), normalize=False).reshape(sp.shape) for x in c]
        if nroots == 1:
            self.converged = bool(conv[0])
            self.eci = float(e[0])
            self.ci = cis[0].cpu().numpy()
        else:
            self.converged = numpy.asarray(conv, dtype=bool)
            self.eci = e
            self.ci = [x.cpu().numpy() for x in cis]
        return self.eci, self.ci


def get_init_guess(norb, nelec, nroots, hdiag_csf, transformer):
    """Determinant-basis guesses from the (packed) CSF diagonal
    (csf.py:51-64)."""
    import torch
    if transformer.ncsf < nroots:
        raise AssertionError("Can't find {} roots among only {} CSFs of symmetry {}".format(
            nroots, transformer.ncsf, transformer.wfnsym))
    hd = hdiag_csf if isinstance(hdiag_csf, torch.Tensor) else torch.as_tensor(
        numpy.asarray(hdiag_csf, dtype=numpy.float64)).cuda()
    if hd.numel() != transformer.ncsf:
        hd = transformer.pack_csf(hd)
    guesses = FCISolver._init_guess_dev(hd, nroots)
    return [transformer.vec_csf2det(g).cpu().numpy() for g in guesses]


def pspace(fci, h1e, eri, norb, nelec, transformer, hdiag_det=None, hdiag_csf=None, npsp=200,
           max_memory=None):
    """(csf_addr, h0) of the npsp lowest-diagonal CSFs (csf.py:246-366)."""
    fci.transformer = transformer
    return fci.pspace(h1e, eri, norb, nelec, hdiag_det=hdiag_det, hdiag_csf=hdiag_csf, npsp=npsp,
                      smult=transformer.smult)


def kernel(fci, h1e, eri, norb, nelec, smult=None, idx_sym=None, ci0=None, transformer=None,
           **kwargs):
    """(e, ci) of ``fci`` in the CSF basis (csf.py:368-497); the point-group
    subspace comes from the solver's orbsym / wfnsym (idx_sym, the
    reference's boolean CSF mask, is accepted only as None)."""
    if idx_sym is not None:
        raise NotImplementedError("pass orbsym / wfnsym to the solver instead of idx_sym")
    if transformer is not None:
        fci.transformer = transformer
    return fci.kernel(h1e, eri, norb, nelec, ci0=ci0, smult=smult, **kwargs)


def csf_solver(mol=None, smult=None, symm=None):
    """CAS CSF solver factory (pyscf.csf_fci.csf_solver, csf_fci/__init__.py:6-12):
    the point-group solver (csf_symm.FCISolver) when ``symm`` is set or,
    unset, when ``mol.symmetry`` is, else the plain CSF solver (which also
    packs by irrep when given ``orbsym``)."""
    if mol is not None and symm is None:
        symm = getattr(mol, "symmetry", None)
    if symm:
        from . import csf_symm
        return csf_symm.FCISolver(mol, smult=smult)
    return FCISolver(mol, smult=smult)

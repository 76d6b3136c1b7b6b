"""Spin ladder operators on CI vectors (pyscf/csf_fci/spin_op.py) on the GPU.

``contract_sladder`` applies S- (op = -1) or S+ (op = +1), changing
(neleca, nelecb) by (op, -op) without changing S; ``mdown`` / ``mup`` rotate
a CI vector between M = S and lower M with the normalisation of the ladder
(norm_sdown / norm_sup).  The contraction is one destination-owned kernel
(fcis_spin_ladder); numpy arrays in give numpy arrays out, device tensors
stay on the device.
"""

import numpy

from . import _lib
from . import direct_spin1 as _ds


def norm_sdown(smult, nelec):
    """|S- |S, M>| for M = (na - nb) / 2 (spin_op.py:12-15)."""
    s2 = smult - 1
    m2 = nelec[0] - nelec[1]
    return numpy.sqrt((s2 - m2 + 2) * (s2 + m2) / 4)


def norm_sup(smult, nelec):
    """|S+ |S, M>| for M = (na - nb) / 2 (spin_op.py:17-20)."""
    s2 = smult - 1
    m2 = nelec[0] - nelec[1]
    return numpy.sqrt((s2 - m2) * (s2 + m2 + 2) / 4)


def contract_sladder(fcivec, norb, nelec, op=-1):
    """S- (op = -1) or S+ (op = +1) times a (na, nb) CI vector, returned as
    the (na + op, nb - op) array; an empty (0, 0) array where the operator
    annihilates every determinant (spin_op.py:22-71)."""
    import torch
    if op not in (-1, 1):
        raise AssertionError("op = -1 or 1")
    na, nb = _ds._unpack_nelec(nelec)
    norb = int(norb)
    if (op == -1 and (na == 0 or nb == norb)) or (op == 1 and (na == norb or nb == 0)):
        return numpy.zeros((0, 0))
    src = _ds.get_space(norb, (na, nb))
    dst = _ds.get_space(norb, (na + op, nb - op))
    on_dev = isinstance(fcivec, torch.Tensor)
    x = fcivec if on_dev else torch.as_tensor(numpy.asarray(fcivec, dtype=numpy.float64))
    x = x.to(device="cuda", dtype=torch.float64).contiguous().reshape(-1)
    if x.numel() != src.ndet:
        raise ValueError("CI vector size %d does not match (%d, %d) in %d orbitals"
                         % (x.numel(), na, nb, norb))
    out = torch.empty(dst.shape, dtype=torch.float64, device="cuda")
    _lib.check(_lib.load().fcis_spin_ladder(src.handle, dst.handle, int(op), _lib.vptr(x),
                                            _lib.vptr(out), _lib.stream_ptr()),
               "fcis_spin_ladder")
    return out if on_dev else out.cpu().numpy()


def contract_sdown(ci, norb, nelec):
    return contract_sladder(ci, norb, nelec, op=-1)


def contract_sup(ci, norb, nelec):
    return contract_sladder(ci, norb, nelec, op=1)


def as_list(ci0):
    if isinstance(ci0, numpy.ndarray) and ci0.ndim == 2:
        return [ci0]
    return list(ci0)


def like_ci0(ci1, ci0):
    if isinstance(ci0, numpy.ndarray) and ci0.ndim == 2:
        return ci1[0]
    return numpy.stack(ci1, axis=0)


def mdown(ci0, norb, nelec, smult):
    """CI vector(s) of M = S (in the (N + smult - 1)/2, (N - smult + 1)/2
    space) rotated to the M of ``nelec`` (spin_op.py:93-121)."""
    ci1 = as_list(ci0)
    na, nb = _ds._unpack_nelec(nelec)
    for i in range(len(ci1)):
        for j in range(((smult - 1) - (na - nb)) // 2, 0, -1):
            ci1[i] = contract_sdown(ci1[i], norb, (na + j, nb - j))
            ci1[i] = ci1[i] / norm_sdown(smult, (na + j, nb - j))
    return like_ci0(ci1, ci0)


def mup(ci0, norb, nelec, smult):
    """CI vector(s) with the M of ``nelec`` rotated to M = S
    (spin_op.py:123-150)."""
    ci1 = as_list(ci0)
    na, nb = _ds._unpack_nelec(nelec)
    for i in range(len(ci1)):
        for j in range(((smult - 1) - (na - nb)) // 2):
            ci1[i] = contract_sup(ci1[i], norb, (na + j, nb - j))
            ci1[i] = ci1[i] / norm_sup(smult, (na + j, nb - j))
    return like_ci0(ci1, ci0)

"""E(h1e, eri) = c^T H c as a torch.autograd.Function.

The forward runs the FP64 sigma on the GPU; the hand-written backward is the
RDM kernel (fcis_trans_rdm12s), the adjoint of the sigma product:

    E      = sum_pq h_pq dm1[p,q] + 1/2 sum_pqrs (pq|rs) dm2[p,q,r,s]
    dE/dh  = dm1                      (pyscf conventions, fci_gas.h:255-268)
    dE/d(pq|rs) = 1/2 dm2[p,q,r,s]

for the symmetric integrals the forward uses (h1e and the 4-fold pair matrix
are symmetrised, (X + X^T)/2, so the gradient is the symmetric part).  With
eri given as a pair matrix eri[PQ, RS] (PQ = max(p,q)(max+1)/2 + min), the
gradient sums the orderings that share a pair index:

    dE/deri[PQ, RS] = 1/2 sum_{(p,q) in PQ} sum_{(r,s) in RS} dm2[p,q,r,s].

dE/dc = 2 H c is returned as well when c requires a gradient.
"""

import numpy

from . import direct_spin1 as _ds


def pair_fold(dm2, norb):
    """sum over index orderings of a (n,n,n,n) array into (npair, npair):
    out[PQ, RS] = sum_{(p,q)->PQ, (r,s)->RS} dm2[p,q,r,s]."""
    n = int(norb)
    p, q = numpy.tril_indices(n)
    offdiag = (p != q).astype(float)
    # fold (r,s)
    a = dm2[:, :, p, q] + offdiag * dm2[:, :, q, p]            # (n, n, npair)
    # fold (p,q)
    return a[p, q, :] + offdiag[:, None] * a[q, p, :]            # (npair, npair)


def energy_gradients(dm1, dm2, norb):
    """(dE/dh1e, dE/deri_pair) from spin-traced RDMs of a real CI vector."""
    gh = 0.5 * (dm1 + dm1.T)
    ge = 0.5 * pair_fold(dm2, norb)
    return gh, 0.5 * (ge + ge.T)


def _torch():
    import torch
    return torch


def _make_function():
    torch = _torch()

    class CIEnergy(torch.autograd.Function):
        @staticmethod
        def forward(ctx, h1e, eri, c, norb, nelec):
            norb = int(norb)
            nelec = _ds._unpack_nelec(nelec)
            sp = _ds.get_space(norb, nelec)
            h1 = h1e.detach().to("cpu", torch.float64).numpy()
            er = _ds.pair_matrix(eri.detach().to("cpu", torch.float64).numpy(), norb)
            h1 = 0.5 * (h1 + h1.T)
            er = 0.5 * (er + er.T)
            h2e = _ds.absorb_h1e(h1, er, norb, nelec, 0.5)
            cd = c.detach().to("cuda", torch.float64).reshape(-1).contiguous()
            if cd.numel() != sp.ndet:
                raise ValueError("CI vector size does not match the determinant count")
            sig = torch.empty_like(cd)
            with _ds.SigmaPlan(sp, h2e, "fp64", _ds._default_workspace()) as plan:
                plan(cd, sig)
            e = torch.dot(cd, sig)
            ctx.save_for_backward(cd, sig)
            ctx.meta = (norb, nelec, h1e.dtype, h1e.device, eri.dtype, eri.device,
                        c.dtype, c.device, tuple(c.shape))
            return e.to(device=h1e.device)

        @staticmethod
        def backward(ctx, ge):
            cd, sig = ctx.saved_tensors
            norb, nelec, hdt, hdev, edt, edev, cdt, cdev, cshape = ctx.meta
            g = float(ge)
            gh = ge_ = gc = None
            if ctx.needs_input_grad[0] or ctx.needs_input_grad[1]:
                dm1, dm2 = _ds.make_rdm12(cd, norb, nelec)
                h_grad, e_grad = energy_gradients(dm1, dm2, norb)
                if ctx.needs_input_grad[0]:
                    gh = torch.as_tensor(g * h_grad, dtype=hdt, device=hdev)
                if ctx.needs_input_grad[1]:
                    ge_ = torch.as_tensor(g * e_grad, dtype=edt, device=edev)
            if ctx.needs_input_grad[2]:
                gc = (2.0 * g * sig).to(dtype=cdt, device=cdev).reshape(cshape)
            return gh, ge_, gc, None, None

    return CIEnergy


_FN = None


def ci_energy(h1e, eri, c, norb, nelec):
    """Differentiable E = c^T H(h1e, eri) c (torch tensors; eri as a pair
    matrix).  FP32 integrals are accepted and promoted to FP64 inside."""
    global _FN
    if _FN is None:
        _FN = _make_function()
    return _FN.apply(h1e, eri, c, norb, nelec)

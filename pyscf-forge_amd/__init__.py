"""MI355X-native direct-CI sigma engine with pyscf's fcisolver API.

    sigma = H c    (pyscf.fci.direct_spin1.contract_2e)

Hand-written gfx950 HIP kernels (libfcisigma.so, C ABI in include/fcisigma.h)
do all determinant-space work: link tables, the t1 gather, the MFMA
contraction, the sign-applying scatter, the Hamiltonian diagonal, RDMs and
the Davidson vector algebra.  This package is the host side: it mirrors the
solver interface the reference's CSF/GAS solvers and CASCI/CASSCF use.
"""
from . import synthetic  # noqa: F401
from . import _lib  # noqa: F401
from .direct_spin1 import (  # noqa: F401
    FCISolver, SigmaPlan, Space, SpinPlan, absorb_h1e, contract_2e, contract_ss, energy,
    gen_linkstr_index,
    get_space, make_hdiag, make_rdm1, make_rdm1s, make_rdm12, make_rdm12s,
    pair_matrix, pspace, release_plans, spin_square_from_rdm12s, trans_rdm1, trans_rdm1s, trans_rdm12,
    trans_rdm12s)

# drop-in alias: code written against pyscf.fci.direct_spin1 can use
# ``from <package> import direct_spin1`` unchanged.
from . import direct_spin1  # noqa: F401,E402
from . import autograd  # noqa: F401,E402
from . import csf_fci  # noqa: F401,E402
from . import batch  # noqa: F401,E402
from . import addons_gas  # noqa: F401,E402
from . import fci_gas  # noqa: F401,E402
from . import gasscf  # noqa: F401,E402
from . import spin_op  # noqa: F401,E402
from . import csf_symm  # noqa: F401,E402

"""Point-group CSF solver (the reference's pyscf/csf_fci/csf_symm.py).

Point groups are the abelian D2h subgroups (irrep ids 0..7, products by
XOR, pyscf's id convention); Dooh and Coov are handled as their D2h / C2v
subgroups, as the reference does (csf_symm.py:42-45).  The solver restricts
the CSF Davidson to the configurations of irrep ``wfnsym`` (csf_fci
packing on the device) and differs from the plain CSF solver the way the
reference's does: ``wfnsym`` may be an irrep name (read with the molecule's
``groupname``), ``orbsym`` / ``wfnsym`` / ``smult`` given to ``kernel``
apply to that call only and the attributes are restored afterwards, the
target irrep defaults to that of ``ci0``'s largest determinant or the
aufbau determinant, and ``get_init_guess`` needs both ``orbsym`` and
``wfnsym``.
"""

import warnings

from . import csf_fci
from .csf_fci import CSFTransformer, ImpossibleSpinError  # noqa: F401

# Irrep ids of the D2h subgroups (pyscf.symm's Cotton ordering): products
# of irreps are XORs of ids.
IRREP_IDS = {
    "D2h": {"Ag": 0, "B1g": 1, "B2g": 2, "B3g": 3, "Au": 4, "B1u": 5, "B2u": 6, "B3u": 7},
    "C2v": {"A1": 0, "A2": 1, "B1": 2, "B2": 3},
    "C2h": {"Ag": 0, "Bg": 1, "Au": 2, "Bu": 3},
    "D2": {"A": 0, "B1": 1, "B2": 2, "B3": 3},
    "Cs": {"A'": 0, 'A"': 1},
    "Ci": {"Ag": 0, "Au": 1},
    "C2": {"A": 0, "B": 1},
    "C1": {"A": 0},
}
# linear groups are solved in these subgroups
LINEAR_SUBGROUP = {"Dooh": "D2h", "Coov": "C2v"}


def irrep_name2id(groupname, name):
    """Irrep id of ``name`` in ``groupname`` (a D2h subgroup, or Dooh/Coov
    read as D2h/C2v); names are case-insensitive."""
    group = LINEAR_SUBGROUP.get(groupname, groupname)
    table = IRREP_IDS.get(group)
    if table is None:
        raise ValueError("point group %r is not an abelian D2h subgroup" % (groupname,))
    folded = {k.lower(): v for k, v in table.items()}
    key = str(name).lower()
    if key not in folded:
        raise ValueError("irrep %r is not in point group %s" % (name, group))
    return folded[key]


def irrep_id2name(groupname, irrep):
    group = LINEAR_SUBGROUP.get(groupname, groupname)
    for k, v in IRREP_IDS[group].items():
        if v == int(irrep):
            return k
    raise ValueError("irrep id %r is not in point group %s" % (irrep, group))


class FCISolver(csf_fci.FCISolver):
    """CSF solver with point-group symmetry (csf_symm.FCISolver,
    csf_symm.py:25-113); ``ci0`` and ``ci`` are determinant-basis arrays."""

    def _groupname(self):
        return getattr(self.mol, "groupname", None)

    def _wfnsym_id(self, wfnsym):
        if wfnsym is None or not isinstance(wfnsym, str):
            return wfnsym
        group = self._groupname()
        if group is None:
            raise ValueError("an irrep name for wfnsym needs mol.groupname")
        return irrep_name2id(group, wfnsym)

    def guess_wfnsym(self, norb, nelec, fcivec=None, orbsym=None, wfnsym=None, **kwargs):
        wfnsym = self.wfnsym if wfnsym is None else wfnsym
        return super().guess_wfnsym(norb, nelec, fcivec, orbsym=orbsym,
                                    wfnsym=self._wfnsym_id(wfnsym))

    def kernel(self, h1e, eri, norb, nelec, ci0=None, **kwargs):
        if self._groupname() in LINEAR_SUBGROUP:
            warnings.warn("Wfn symmetry for Dooh/Coov not supported. Wfn symmetry is mapped "
                          "to D2h/C2v group.")
        saved = (self.orbsym, self.wfnsym)
        orbsym = kwargs.pop("orbsym", None)
        wfnsym = kwargs.pop("wfnsym", None)
        smult = kwargs.pop("smult", None)
        if smult is not None:
            self.smult = smult
        try:
            if orbsym is not None:
                self.orbsym = orbsym
            target = self.guess_wfnsym(norb, nelec, ci0, wfnsym=wfnsym)
            if target is not None and target > 9:
                raise NotImplementedError("High-momentum point groups in Dooh/Coov")
            return super().kernel(h1e, eri, norb, nelec, ci0=ci0, wfnsym=target, **kwargs)
        finally:
            self.orbsym, self.wfnsym = saved

    def get_init_guess(self, norb, nelec, nroots, hdiag_csf, **kwargs):
        """Determinant-basis guesses of irrep ``wfnsym`` (kwargs override the
        attributes and are kept, as in the reference)."""
        self.orbsym = kwargs.get("orbsym", self.orbsym)
        self.wfnsym = kwargs.get("wfnsym", self.wfnsym)
        if self.orbsym is None or self.wfnsym is None:
            raise AssertionError("get_init_guess of the point-group solver needs orbsym and "
                                 "wfnsym")
        self.wfnsym = self._wfnsym_id(self.wfnsym)
        self.norb, self.nelec = int(norb), self._nelec(nelec)
        t = self.check_transformer_cache()
        return csf_fci.get_init_guess(norb, nelec, nroots, hdiag_csf, t)

"""Point-group CSF solver (pyscf/csf_fci/csf_symm.py).

The CSF solver of csf_fci handles point-group symmetry itself: with
``orbsym`` (D2h-subgroup irrep ids) it packs CSF vectors to the
configurations of irrep ``wfnsym`` (guessed from the aufbau determinant
when unset), so this module's FCISolver is that class.
"""

from .csf_fci import FCISolver, CSFTransformer, ImpossibleSpinError  # noqa: F401

"""CPU restatement of the restricted (GAS) determinant space and its sigma.

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the CPU
baseline of bench.py as the checker, never by the product package.

What is restated (numpy, small cases):
  * the sector pool: every alpha and beta occupation tuple of the block set
    D plus their single inter-space excitations, unique and sorted
    lexicographically (pyscf/lib/mcscf/fci_string_gas.c:636-692), which
    fixes the sector ids;
  * sector string addressing: per-space combinatorial addresses combined
    row-major with the last space fastest (init_sectors_from_pool
    fci_string_gas.c:694-732, gas_str2addr_sector / gas_addr2str_sector
    :462-489);
  * the block layout: blocks in canonical order, each an alpha-major
    rectangle at offset = running determinant count (fill_blocks_from_raw
    :759-791, gas_det_addr :497-503);
  * gas2fci / fci2gas (pyscf/mcscf/fci_gas.py:141-208);
  * the restricted Hamiltonian as the projection of the CAS one,
    sigma_GAS = fci2gas(H gas2fci(c)), which is how the reference's own test
    checks libfci_gas (pyscf/mcscf/test/test_gasci.py:522-575); the same
    projection gives the diagonal, and the RDMs of a GAS vector are those of
    its embedding.

Pinning: tests/test_gas_oracle.py compares every function here with the
reference's libfci_gas (compiled from its sources into oracle/_ref/) on
multi-space, multi-block cases and with the fixtures it wrote under
tests/golden/, and the restricted eigenvalues with a dense projected H.
"""

import functools
import itertools

import numpy

from . import contract_2e, make_hdiag, trans_rdm12s, num_strings


@functools.lru_cache(maxsize=None)
def _local_strings(norb, nocc):
    """Bit strings of nocc electrons in norb orbitals in combinatorial
    (colex) address order."""
    out = []
    for occ in itertools.combinations(range(norb), nocc):
        out.append(sum(1 << p for p in occ))
    return numpy.asarray(sorted(out, key=_colex_key(nocc)), dtype=numpy.int64)


def _colex_key(nocc):
    def key(s):
        occ = [p for p in range(64) if (s >> p) & 1]
        return sum(_comb(p, i + 1) for i, p in enumerate(occ))
    return key


def _comb(n, k):
    if k < 0 or k > n:
        return 0
    r = 1
    for i in range(1, k + 1):
        r = r * (n - k + i) // i
    return r


def colex_addr(s):
    """CAS string address (combinatorial number system)."""
    occ = [p for p in range(64) if (s >> p) & 1]
    return sum(_comb(p, i + 1) for i, p in enumerate(occ))


class GasLayout:
    """Sectors, blocks and strings of a GAS space (norb per space, block set
    D as rows (alpha occ..., beta occ...) in canonical order)."""

    def __init__(self, gas_orbs, nelec, blocks):
        self.norb = tuple(int(x) for x in gas_orbs)
        self.ngas = len(self.norb)
        self.start = tuple(int(x) for x in numpy.cumsum((0,) + self.norb)[:-1])
        self.norb_total = sum(self.norb)
        self.nelec = (int(nelec[0]), int(nelec[1]))
        blocks = numpy.asarray(blocks, dtype=numpy.int64).reshape(-1, 2 * self.ngas)
        rows = [tuple(int(v) for v in r) for r in blocks]
        if rows != sorted(set(rows)):
            raise ValueError("blocks must be unique and sorted")
        self.blocks = rows
        pool = set()
        for r in rows:
            for occ in (r[:self.ngas], r[self.ngas:]):
                pool.add(occ)
                for a in range(self.ngas):
                    for b in range(self.ngas):
                        if a != b and occ[a] > 0 and occ[b] < self.norb[b]:
                            o = list(occ)
                            o[a] -= 1
                            o[b] += 1
                            pool.add(tuple(o))
        self.sectors = sorted(pool)
        self.sector_id = {s: i for i, s in enumerate(self.sectors)}
        self.sector_nstr = [int(numpy.prod([_comb(n, k) for n, k in zip(self.norb, s)]))
                            for s in self.sectors]
        self.block_desc = []
        off = 0
        for r in rows:
            sa = self.sector_id[r[:self.ngas]]
            sb = self.sector_id[r[self.ngas:]]
            na, nb = self.sector_nstr[sa], self.sector_nstr[sb]
            self.block_desc.append(dict(offset=off, sa=sa, sb=sb, na=na, nb=nb))
            off += na * nb
        self.ndet = off
        self._strings = {}

    def sector_strings(self, sid):
        """All strings of a sector in sector-address order."""
        if sid not in self._strings:
            occ = self.sectors[sid]
            per = [_local_strings(n, k) << st
                   for n, k, st in zip(self.norb, occ, self.start)]
            # row-major over spaces, the last space fastest
            out = numpy.zeros(1, dtype=numpy.int64)
            for x in per:
                out = (out[:, None] + x[None, :]).reshape(-1)
            self._strings[sid] = out
        return self._strings[sid]

    def det_strings(self):
        """(alpha strings, beta strings) of every GAS determinant, in the
        space's order (uint64)."""
        sa = numpy.empty(self.ndet, dtype=numpy.uint64)
        sb = numpy.empty(self.ndet, dtype=numpy.uint64)
        for d in self.block_desc:
            a = self.sector_strings(d["sa"]).astype(numpy.uint64)
            b = self.sector_strings(d["sb"]).astype(numpy.uint64)
            o, na, nb = d["offset"], d["na"], d["nb"]
            sa[o:o + na * nb] = numpy.repeat(a, nb)
            sb[o:o + na * nb] = numpy.tile(b, na)
        return sa, sb

    def cas_index(self):
        """(ndet,) flat (nstr_alpha, nstr_beta) C-order CAS index of every
        GAS determinant."""
        nstrb = num_strings(self.norb_total, self.nelec[1])
        idx = numpy.empty(self.ndet, dtype=numpy.int64)
        for d in self.block_desc:
            ia = numpy.array([colex_addr(int(s)) for s in self.sector_strings(d["sa"])])
            ib = numpy.array([colex_addr(int(s)) for s in self.sector_strings(d["sb"])])
            idx[d["offset"]:d["offset"] + d["na"] * d["nb"]] = (
                ia[:, None] * nstrb + ib[None, :]).reshape(-1)
        return idx

    @property
    def cas_shape(self):
        return (num_strings(self.norb_total, self.nelec[0]),
                num_strings(self.norb_total, self.nelec[1]))


def gas2fci(c, layout):
    """Embed a GAS vector in the CAS (nstr_alpha, nstr_beta) array
    (fci_gas.py:141-182)."""
    c = numpy.asarray(c, dtype=numpy.float64).reshape(-1)
    if c.size != layout.ndet:
        raise ValueError("CI vector size does not match GAS determinant count")
    out = numpy.zeros(layout.cas_shape)
    out.reshape(-1)[layout.cas_index()] = c
    return out


def fci2gas(c, layout):
    """Project a CAS array onto the GAS determinants (fci_gas.py:185-208)."""
    c = numpy.asarray(c, dtype=numpy.float64).reshape(-1)
    return c[layout.cas_index()].copy()


def gas_contract_2e(h2e, c, layout):
    """Restricted sigma: fci2gas(H gas2fci(c))."""
    full = gas2fci(c, layout)
    return fci2gas(contract_2e(h2e, full, layout.norb_total, layout.nelec), layout)


def gas_make_hdiag(h1e, eri, layout):
    return fci2gas(make_hdiag(h1e, eri, layout.norb_total, layout.nelec), layout)


def gas_trans_rdm12s(bra, ket, layout):
    return trans_rdm12s(gas2fci(bra, layout), gas2fci(ket, layout),
                        layout.norb_total, layout.nelec)


def gas_hamiltonian(h2e, layout):
    """Dense restricted H (small cases): columns are sigma of unit vectors."""
    n = layout.ndet
    H = numpy.empty((n, n))
    e = numpy.zeros(n)
    for j in range(n):
        e[j] = 1.0
        H[:, j] = gas_contract_2e(h2e, e, layout)
        e[j] = 0.0
    return 0.5 * (H + H.T)

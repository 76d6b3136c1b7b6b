"""GAS restriction front end: every public restriction form is expanded to
the canonical block set D, rows (alpha occupations..., beta occupations...)
per GAS subspace, unique and lexicographically sorted
(pyscf/mcscf/addons_gas.py: normalize_gas_spec :653-705, the four
restriction types :645-650, blocks_from_supergroups :360-428,
supergroups_from_cumulative_occ :431-524, the RAS mapping :562-608,
is_spin_complete :782-798, check_kernel_limits :808-861).

Names, argument meaning and error types follow the reference: non-integer
occupations raise TypeError, inconsistent ones ValueError, an unknown
restriction type NotImplementedError; messages end with the usage text.
"""

from collections.abc import Mapping
from functools import lru_cache
import itertools
import math

import numpy

from .direct_spin1 import _unpack_nelec

GAS_MAX_ORB = 63
GAS_MAX_LOCAL_ORB = 31
GAS_MAX_NGAS = 15
GAS_MAX_SECTOR = 65534
GAS_MAX_DET = 2**32 - 1
GAS_MAX_BLOCK = 2**31 - 1
GAS_RESTR_SPIN_SUPERGROUP = "spin-supergroup"
GAS_RESTR_SUPERGROUP = "supergroup"
GAS_RESTR_CUMULATIVE_OCC = "cumulative-occ"
GAS_RESTR_RAS = "ras"

_USAGE = {
    GAS_RESTR_SPIN_SUPERGROUP:
        "gas_restr_type='spin-supergroup': gas_restr is a non-empty array of shape "
        "(nrow, 2*ngas); each row is (alpha occupations, beta occupations).",
    GAS_RESTR_SUPERGROUP:
        "gas_restr_type='supergroup': gas_restr is a non-empty array of shape "
        "(nrow, ngas); each row contains total occupations.",
    GAS_RESTR_CUMULATIVE_OCC:
        "gas_restr_type='cumulative-occ': gas_restr has shape (ngas, 2), with "
        "non-decreasing [minimum, maximum] cumulative occupations and a final bound "
        "equal to the active electrons.",
    GAS_RESTR_RAS:
        "gas_restr_type='ras': gas_orbs has exactly three entries (RAS1, RAS2, RAS3), "
        "which may be zero, and gas_restr is exactly {'max_holes': integer, "
        "'max_particles': integer}.",
}


def gas_restr_usage(gas_restr_type=None):
    """The input contract of the GAS restriction forms."""
    head = ("GAS restriction input rules (only row order and duplicate rows are "
            "normalized):")
    tail = ("All occupations, bounds, orbital counts, and orbital indices must be "
            "integers; floating-point and Boolean values are rejected.")
    body = [_USAGE[gas_restr_type]] if gas_restr_type in _USAGE else list(_USAGE.values())
    return "\n".join([head] + body + [tail])


def _is_int(v):
    return (isinstance(v, (int, numpy.integer))
            and not isinstance(v, (bool, numpy.bool_)))


def _int_array(values, label):
    obj = numpy.asarray(values, dtype=object)
    if not all(_is_int(v) for v in obj.reshape(-1)):
        raise TypeError("%s must contain integers" % label)
    arr = numpy.asarray(values).astype(numpy.int64) if obj.size else numpy.zeros(obj.shape, numpy.int64)
    if arr.size and (arr.min() < -2**31 or arr.max() > 2**31 - 1):
        raise ValueError("%s values exceed int32 range" % label)
    return arr.astype(numpy.int32)


def _int_vector(values, label, length=None):
    arr = _int_array(values, label)
    if arr.ndim != 1:
        raise ValueError("%s must be a one-dimensional sequence" % label)
    if length is not None and arr.size != length:
        raise ValueError("%s must contain exactly %d entries" % (label, length))
    if arr.size == 0:
        raise ValueError("%s must not be empty" % label)
    return tuple(int(v) for v in arr)


def _rows(values, ncol, label):
    arr = _int_array(values, label)
    if arr.ndim != 2 or arr.shape[1] != ncol:
        raise ValueError("%s must have shape (nrow, %d)" % (label, ncol))
    if arr.shape[0] == 0:
        raise ValueError("%s must contain at least one row" % label)
    return numpy.ascontiguousarray(numpy.unique(arr, axis=0), dtype=numpy.int32)


def normalize_blocks(blocks, ngas):
    """Canonical spin-supergroup set D: unique rows, sorted."""
    return _rows(blocks, 2 * int(ngas), "spin-supergroups")


def normalize_supergroups(supergroups, ngas):
    """Canonical spin-free supergroup set G."""
    return _rows(supergroups, int(ngas), "supergroups")


def normalize_cumulative_occ(bounds, ngas):
    """Validated (ngas, 2) cumulative [min, max] bounds (order kept)."""
    arr = _int_array(bounds, "cumulative occupation bounds")
    if arr.shape != (int(ngas), 2):
        raise ValueError("cumulative occupation bounds must have shape (%d, 2)" % int(ngas))
    return numpy.ascontiguousarray(arr, dtype=numpy.int32)


def _check_orbs(gas_orbs, allow_empty=False):
    if not 0 < len(gas_orbs) <= GAS_MAX_NGAS:
        raise ValueError("number of GAS spaces exceeds C kernel limit")
    lo = 0 if allow_empty else 1
    if any(n < lo or n > GAS_MAX_LOCAL_ORB for n in gas_orbs):
        raise ValueError("each %s space must have %d..31 orbitals"
                         % ("RAS" if allow_empty else "GAS", lo))
    if sum(gas_orbs) <= 0:
        raise ValueError("at least one active orbital is required")
    if sum(gas_orbs) > GAS_MAX_ORB:
        raise ValueError("total active orbitals exceed 63")


def _alpha_splits(gas_orbs, total, na):
    """All alpha occupations a (per space) with sum(a) = na and the beta
    remainder total - a within each space's capacity, in lexicographic order."""
    ranges = [range(max(0, t - n), min(n, t) + 1) for n, t in zip(gas_orbs, total)]
    for a in itertools.product(*ranges):
        if sum(a) == na:
            yield a


def blocks_from_supergroups(gas_orbs, nelec, supergroups):
    """Expand spin-free supergroups G into the canonical block set D."""
    gas_orbs = _int_vector(gas_orbs, "gas_orbs")
    _check_orbs(gas_orbs)
    na, nb = _unpack_nelec(nelec)
    if na < 0 or nb < 0 or max(na, nb) > sum(gas_orbs):
        raise ValueError("invalid alpha/beta electron count")
    G = normalize_supergroups(supergroups, len(gas_orbs))
    if (G < 0).any():
        raise ValueError("supergroup occupations must be non-negative")
    if (G > 2 * numpy.asarray(gas_orbs)).any():
        raise ValueError("supergroup occupation exceeds local GAS capacity")
    if (G.sum(axis=1) != na + nb).any():
        raise ValueError("every supergroup must sum to the requested electron count")
    rows = []
    for g in G.tolist():
        for a in _alpha_splits(gas_orbs, g, na):
            rows.append(tuple(a) + tuple(t - x for t, x in zip(g, a)))
    if not rows:
        raise ValueError("supergroups generate no legal spin-supergroups")
    if len(rows) > GAS_MAX_BLOCK:
        raise ValueError("number of blocks exceeds C int32 input limit")
    D = normalize_blocks(rows, len(gas_orbs))
    check_kernel_limits(gas_orbs, (na, nb), D)
    return D


def supergroups_from_cumulative_occ(gas_orbs, nelec, bounds, allow_empty=False):
    """Spin-free supergroups allowed by cumulative occupation bounds: the
    electrons in spaces 0..k lie in [bounds[k,0], bounds[k,1]]."""
    gas_orbs = _int_vector(gas_orbs, "gas_orbs")
    _check_orbs(gas_orbs, allow_empty)
    na, nb = _unpack_nelec(nelec)
    ne = na + nb
    if na < 0 or nb < 0 or max(na, nb) > sum(gas_orbs):
        raise ValueError("invalid alpha/beta electron count")
    B = normalize_cumulative_occ(bounds, len(gas_orbs))
    lo, hi = B[:, 0], B[:, 1]
    cap = 2 * numpy.cumsum(gas_orbs)
    if (lo < 0).any():
        raise ValueError("cumulative occupations must be non-negative")
    if (lo > hi).any():
        raise ValueError("cumulative minimum exceeds maximum")
    if (numpy.diff(lo) < 0).any() or (numpy.diff(hi) < 0).any():
        raise ValueError("cumulative bounds must be non-decreasing")
    if (hi > cap).any():
        raise ValueError("cumulative occupation exceeds orbital capacity")
    if (hi > ne).any():
        raise ValueError("cumulative occupation exceeds active electrons")
    if lo[-1] != ne or hi[-1] != ne:
        raise ValueError("the final cumulative bound must equal the active electron count")
    ng = len(gas_orbs)

    @lru_cache(maxsize=None)
    def tails(k, acc):
        """Occupation tuples of spaces k.. given acc electrons before k."""
        if k == ng:
            return ((),) if acc == ne else ()
        out = []
        for x in range(0, 2 * gas_orbs[k] + 1):
            c = acc + x
            if lo[k] <= c <= hi[k]:
                out.extend((x,) + t for t in tails(k + 1, c))
        return tuple(out)

    rows = tails(0, 0)
    if not rows:
        raise ValueError("cumulative bounds generate no legal supergroups")
    if len(rows) > GAS_MAX_BLOCK:
        raise ValueError("number of generated supergroups exceeds int32 limit")
    return normalize_supergroups(rows, ng)


def blocks_from_cumulative_occ(gas_orbs, nelec, bounds):
    return blocks_from_supergroups(
        gas_orbs, nelec, supergroups_from_cumulative_occ(gas_orbs, nelec, bounds))


def _ras_limits(gas_restr):
    if not isinstance(gas_restr, Mapping):
        raise TypeError("ras gas_restr must be a mapping with max_holes and max_particles")
    need = {"max_holes", "max_particles"}
    if need - set(gas_restr):
        raise ValueError("missing RAS restriction keys: %s" % ", ".join(sorted(need - set(gas_restr))))
    if set(gas_restr) - need:
        raise ValueError("unknown RAS restriction keys: %s" % ", ".join(sorted(set(gas_restr) - need)))
    h, p = gas_restr["max_holes"], gas_restr["max_particles"]
    if not _is_int(h):
        raise TypeError("max_holes must be an integer")
    if not _is_int(p):
        raise TypeError("max_particles must be an integer")
    if h < 0 or p < 0:
        raise ValueError("RAS hole and particle limits must be non-negative")
    return int(h), int(p)


def _spec_ras(gas_orbs, nelec, gas_restr):
    """RAS1/RAS2/RAS3 with at most max_holes holes in RAS1 and max_particles
    electrons in RAS3, as cumulative bounds; empty spaces are dropped from
    the kernel space (addons_gas.py:562-608)."""
    orbs = _int_vector(gas_orbs, "ras gas_orbs", length=3)
    if any(n < 0 or n > GAS_MAX_LOCAL_ORB for n in orbs):
        raise ValueError("each RAS space must have 0..31 orbitals")
    if sum(orbs) <= 0:
        raise ValueError("at least one RAS orbital is required")
    if sum(orbs) > GAS_MAX_ORB:
        raise ValueError("total active orbitals exceed 63")
    holes, parts = _ras_limits(gas_restr)
    n1, n2, n3 = orbs
    if holes > 2 * n1:
        raise ValueError("max_holes exceeds the RAS1 electron capacity")
    if parts > 2 * n3:
        raise ValueError("max_particles exceeds the RAS3 electron capacity")
    ne = sum(_unpack_nelec(nelec))
    first = [max(0, 2 * n1 - holes), min(2 * n1, ne)]
    second = [max(0, ne - parts, first[0]), min(2 * (n1 + n2), ne)]
    bounds = numpy.asarray([first, second, [ne, ne]], dtype=numpy.int32)
    G = supergroups_from_cumulative_occ(orbs, nelec, bounds, allow_empty=True)
    keep = numpy.asarray(orbs) > 0
    korbs = tuple(int(n) for n, k in zip(orbs, keep) if k)
    KG = normalize_supergroups(G[:, keep], len(korbs))
    D = blocks_from_supergroups(korbs, nelec, KG)
    return korbs, D, {"ras_orbs": orbs, "max_holes": holes, "max_particles": parts,
                      "cumulative_bounds": bounds, "canonical_supergroups": KG,
                      "conceptual_supergroups": G,
                      "empty_spaces_removed": tuple(i + 1 for i, n in enumerate(orbs)
                                                    if n == 0)}


def _spec_spin_supergroup(gas_orbs, nelec, gas_restr):
    D = normalize_blocks(gas_restr, len(gas_orbs))
    check_kernel_limits(gas_orbs, nelec, D)
    return gas_orbs, D, {"canonical_spin_supergroups": D}


def _spec_supergroup(gas_orbs, nelec, gas_restr):
    G = normalize_supergroups(gas_restr, len(gas_orbs))
    D = blocks_from_supergroups(gas_orbs, nelec, G)
    return gas_orbs, D, {"canonical_supergroups": G, "canonical_spin_supergroups": D}


def _spec_cumulative(gas_orbs, nelec, gas_restr):
    bounds = normalize_cumulative_occ(gas_restr, len(gas_orbs))
    G = supergroups_from_cumulative_occ(gas_orbs, nelec, bounds)
    D = blocks_from_supergroups(gas_orbs, nelec, G)
    return gas_orbs, D, {"cumulative_bounds": bounds, "canonical_supergroups": G,
                         "canonical_spin_supergroups": D}


_SPECS = {
    GAS_RESTR_SPIN_SUPERGROUP: _spec_spin_supergroup,
    GAS_RESTR_SUPERGROUP: _spec_supergroup,
    GAS_RESTR_CUMULATIVE_OCC: _spec_cumulative,
    GAS_RESTR_RAS: _spec_ras,
}


def normalize_gas_spec(gas_orbs, nelec, gas_restr=None,
                       gas_restr_type=GAS_RESTR_SPIN_SUPERGROUP, return_info=False):
    """(kernel gas_orbs, canonical D[, info]) for one public GAS restriction."""
    try:
        if not isinstance(gas_restr_type, str):
            raise TypeError("gas_restr_type must be a string")
        if gas_restr_type not in _SPECS:
            raise NotImplementedError("unsupported gas_restr_type %r; supported types: %s"
                                      % (gas_restr_type, ", ".join(sorted(_SPECS))))
        orbs = _int_vector(gas_orbs, "gas_orbs",
                           length=3 if gas_restr_type == GAS_RESTR_RAS else None)
        info = {"gas_restr_type": gas_restr_type, "user_gas_orbs": orbs}
        if gas_restr is None:
            if gas_restr_type != GAS_RESTR_SPIN_SUPERGROUP:
                raise ValueError("gas_restr is required for gas_restr_type %r"
                                 % gas_restr_type)
            if len(orbs) != 1:
                raise ValueError("gas_restr is required for multi-space GAS")
            D = numpy.asarray([_unpack_nelec(nelec)], dtype=numpy.int32)
            check_kernel_limits(orbs, nelec, D)
            info["canonical_spin_supergroups"] = D
        else:
            orbs, D, extra = _SPECS[gas_restr_type](orbs, nelec, gas_restr)
            info.update(extra)
        info["kernel_gas_orbs"] = orbs
    except (TypeError, ValueError, NotImplementedError) as err:
        raise type(err)("%s\n%s" % (err, gas_restr_usage(
            gas_restr_type if isinstance(gas_restr_type, str) else None))) from err
    return (orbs, D, info) if return_info else (orbs, D)


def normalize_gas_restr(gas_orbs, nelec, gas_restr=None,
                        gas_restr_type=GAS_RESTR_SPIN_SUPERGROUP):
    """Canonical block set D of one public restriction."""
    return normalize_gas_spec(gas_orbs, nelec, gas_restr, gas_restr_type)[1]


def is_spin_complete(gas_orbs, nelec, blocks):
    """True when D holds every alpha/beta split of each of its spin-free
    supergroups (the space is then invariant under S^2)."""
    gas_orbs = _int_vector(gas_orbs, "gas_orbs")
    D = normalize_blocks(blocks, len(gas_orbs))
    ng = len(gas_orbs)
    G = numpy.unique(D[:, :ng] + D[:, ng:], axis=0)
    full = blocks_from_supergroups(gas_orbs, nelec, G)
    return {tuple(r) for r in D.tolist()} == {tuple(r) for r in full.tolist()}


def _nstr(norb, occ):
    return math.prod(math.comb(int(n), int(k)) for n, k in zip(norb, occ))


def check_kernel_limits(norb, nelec, blocks):
    """Validate D against the space limits; return its dimensions."""
    norb = _int_vector(norb, "gas_orbs")
    _check_orbs(norb)
    na, nb = _unpack_nelec(nelec)
    if na < 0 or nb < 0 or max(na, nb) > sum(norb):
        raise ValueError("invalid alpha/beta electron count")
    ng = len(norb)
    D = normalize_blocks(blocks, ng)
    if D.shape[0] > GAS_MAX_BLOCK:
        raise ValueError("number of blocks exceeds C int32 input limit")
    A, B = D[:, :ng], D[:, ng:]
    if (D < 0).any():
        raise ValueError("block occupations must be non-negative")
    if (A > numpy.asarray(norb)).any():
        raise ValueError("alpha block occupation exceeds local GAS orbitals")
    if (B > numpy.asarray(norb)).any():
        raise ValueError("beta block occupation exceeds local GAS orbitals")
    if (A.sum(axis=1) != na).any():
        raise ValueError("every alpha block must sum to the requested nelec")
    if (B.sum(axis=1) != nb).any():
        raise ValueError("every beta block must sum to the requested nelec")
    sectors = {tuple(r) for r in A.tolist()} | {tuple(r) for r in B.tolist()}
    if len(sectors) > GAS_MAX_SECTOR:
        raise ValueError("number of GAS sectors exceeds uint16_t kernel limit")
    if any(_nstr(norb, s) > GAS_MAX_DET for s in sectors):
        raise ValueError("sector string count exceeds uint32_t kernel limit")
    ndet = 0
    for a, b in zip(A.tolist(), B.tolist()):
        ndet += _nstr(norb, a) * _nstr(norb, b)
        if ndet > GAS_MAX_DET:
            raise ValueError("determinant count exceeds uint32_t kernel limit")
    return {"ngas": ng, "norb_total": sum(norb), "na": na, "nb": nb,
            "nblock": int(D.shape[0]), "nsector_upper": len(sectors),
            "ndet_estimate": int(ndet)}


def pair_index(p, q):
    hi, lo = max(int(p), int(q)), min(int(p), int(q))
    return hi * (hi + 1) // 2 + lo

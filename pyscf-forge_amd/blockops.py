"""Block vector algebra over many device vectors of different lengths, one
kernel launch per call (fcis_gram_list / fcis_gemm_list / fcis_precond_list):
the Davidson work of the single solver and of the batched solver."""

import numpy

from . import _lib

BLOCK_DTYPE = numpy.dtype([("n", "<i8"), ("nx", "<i4"), ("ny", "<i4"), ("xoff", "<i4"),
                         ("yoff", "<i4"), ("ooff", "<i4"), ("pad", "<i4"), ("beta", "<f8")])
PRE_DTYPE = numpy.dtype([("hdiag", "<u8"), ("r", "<u8"), ("out", "<u8"), ("n", "<i8"),
                         ("e", "<f8")])


class BlockOps:
    """Block vector algebra over many device vectors of different lengths,
    one kernel launch per call (fcis_gram_list / fcis_gemm_list /
    fcis_precond_list).  A block is an array of device addresses."""

    def __init__(self):
        self.lib = _lib.load()

    @staticmethod
    def _entries(blocks, nout_fn):
        ent = numpy.zeros(len(blocks), BLOCK_DTYPE)
        xo = yo = oo = 0
        for i, b in enumerate(blocks):
            nx, ny = len(b[0]), len(b[1])
            ent[i] = (b[2], nx, ny, xo, yo, oo, 0, b[4] if len(b) > 4 else 0.0)
            xo += nx
            yo += ny
            oo += nout_fn(nx, ny)
        xs = numpy.concatenate([numpy.asarray(b[0], dtype=numpy.uint64) for b in blocks])
        ys = numpy.concatenate([numpy.asarray(b[1], dtype=numpy.uint64) for b in blocks])
        return ent, xs, ys, oo

    def gram(self, blocks):
        """blocks: list of (x_addrs, y_addrs, n) -> list of (nx, ny) X^T Y."""
        if not blocks:
            return []
        ent, xs, ys, nout = self._entries(blocks, lambda nx, ny: nx * ny)
        out = numpy.empty(max(nout, 1))
        _lib.check(self.lib.fcis_gram_list(len(blocks), ent.ctypes.data, len(xs), xs.ctypes.data,
                                           len(ys), ys.ctypes.data, nout, out.ctypes.data,
                                           _lib.stream_ptr()), "fcis_gram_list")
        res, o = [], 0
        for b in blocks:
            nx, ny = len(b[0]), len(b[1])
            res.append(out[o:o + nx * ny].reshape(nx, ny))
            o += nx * ny
        return res

    def gemm(self, blocks):
        """blocks: list of (x_addrs, y_addrs, n, coef (nx, ny), beta):
        Y = beta Y + X coef."""
        if not blocks:
            return
        ent, xs, ys, ncoef = self._entries(blocks, lambda nx, ny: nx * ny)
        cf = numpy.concatenate([numpy.ascontiguousarray(b[3], dtype=numpy.float64).reshape(-1)
                                for b in blocks] + [numpy.zeros(1)])
        _lib.check(self.lib.fcis_gemm_list(len(blocks), ent.ctypes.data, len(xs), xs.ctypes.data,
                                           len(ys), ys.ctypes.data, ncoef, cf.ctypes.data,
                                           _lib.stream_ptr()), "fcis_gemm_list")

    def precond(self, entries, level_shift):
        """entries: list of (hdiag_addr, r_addr, out_addr, n, e)."""
        if not entries:
            return
        ent = numpy.array(entries, dtype=PRE_DTYPE)
        _lib.check(self.lib.fcis_precond_list(len(entries), ent.ctypes.data, float(level_shift),
                                              _lib.stream_ptr()), "fcis_precond_list")

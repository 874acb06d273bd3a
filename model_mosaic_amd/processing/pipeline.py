"""StreamPipe: host-staged batches through the device with copies and compute overlapped.

The ``*_large`` operators move host arrays through the device kernels in
batches (z slabs, frame ranges).  ``StreamPipe`` runs such a loop on three
HIP streams with two ping-pong slots per side (the pattern of the
reference's processing/pipeline.py, built here on torch's ROCm streams and
page-locked host buffers):

    load(k, host_in)        CPU: caller's batch k -> pinned slot (main thread)
    H2D                     pinned slot -> device slot           (stream h2d)
    compute(k, din, dout)   device slot -> device slot           (stream comp)
    D2H                     device slot -> pinned slot           (stream d2h)
    store(k, host_out)      CPU: pinned slot -> caller's batch k (main thread)

Batch k uploads while k - 1 computes and k - 2 downloads; the CPU copies of
``load`` and ``store`` overlap the device work of the neighbouring batches.
Events order the slots: an input slot is overwritten only after the compute
that read it, an output slot only after the download that read it.  All
streams keep the default priority.
"""
from __future__ import annotations

import numpy as np
import torch

from model_mosaic_amd._carriers import pinned_empty

# Largest page-locked block one BandedPinned band asks for (processing/pipeline.py:60-69 keeps each
# host allocation under a per-call cap of this size).
MAX_PINNED_BAND_BYTES = 32 << 30


def pick_n_bands(shape, dtype, band_axis: int = 1, min_bands: int = 1,
                 max_band_bytes: int = MAX_PINNED_BAND_BYTES) -> int:
    """Smallest divisor of shape[band_axis] that is >= min_bands and keeps every band within max_band_bytes
    (processing/pipeline.py:71-89 gives the rule); one row per band if no divisor does."""
    rows = int(shape[band_axis])
    total = int(np.prod(shape)) * np.dtype(dtype).itemsize
    need = max(int(min_bands), -(-total // int(max_band_bytes)))
    return next((k for k in range(need, rows + 1) if rows % k == 0), rows)


class BandedPinned:
    """A host array of ``shape`` kept as ``n_bands`` page-locked blocks along ``band_axis``.

    Same interface as processing/pipeline.py:92-252 (the buffers the large operators hand out and take back):
    ``shape``, ``dtype``, ``ndim``, ``nbytes``, ``band_axis``, ``n_bands``, ``band_rows``, ``bands`` (numpy views
    of the blocks, filled band by band by callers such as step3_propagation_large.py:210-226), ``fill``,
    ``self[key]`` / ``self[key] = value`` for integer and unit-step slice keys (a read inside one band is a
    view, across bands a new array), ``copy_to`` / ``copy_from`` (band-wise, no full-size temporary),
    ``copy_to_gpu`` (torch device destination) and ``__array__`` (materialises the whole array)."""

    def __init__(self, shape, dtype, n_bands: int, band_axis: int = 1, pinned: bool = True):
        self.shape = tuple(int(v) for v in shape)
        self.dtype = np.dtype(dtype)
        self.ndim = len(self.shape)
        self.band_axis = int(band_axis)
        rows = self.shape[self.band_axis]
        if n_bands < 1 or rows % n_bands:
            raise ValueError(f"BandedPinned: shape[{self.band_axis}]={rows} not divisible by n_bands={n_bands}")
        self.n_bands = int(n_bands)
        self.band_rows = rows // self.n_bands
        bshape = list(self.shape)
        bshape[self.band_axis] = self.band_rows
        tdtype = torch.from_numpy(np.zeros(0, self.dtype)).dtype
        # pinned=False keeps the blocks in pageable memory (host-side use without a GPU runtime)
        make = pinned_empty if pinned else (lambda shp, dt: torch.empty(shp, dtype=dt))
        self._blocks = [make(tuple(bshape), tdtype) for _ in range(self.n_bands)]
        self.bands = [b.numpy() for b in self._blocks]

    @property
    def nbytes(self) -> int:
        return int(np.prod(self.shape)) * self.dtype.itemsize

    def __array__(self, dtype=None, copy=None):
        whole = np.concatenate(self.bands, axis=self.band_axis)
        return whole if dtype is None else whole.astype(dtype)

    def fill(self, value):
        for b in self.bands:
            b.fill(value)

    # key handling: a tuple padded with full slices; the band-axis entry is an int or a unit-step slice
    def _key(self, key):
        key = key if isinstance(key, tuple) else (key,)
        if len(key) > self.ndim:
            raise IndexError(f"BandedPinned: {len(key)} indices for {self.ndim} axes")
        return list(key) + [slice(None)] * (self.ndim - len(key))

    def _span(self, k):
        rows = self.shape[self.band_axis]
        if isinstance(k, slice):
            lo, hi, step = k.indices(rows)
            if step != 1:
                raise NotImplementedError("BandedPinned: only unit-step slices along the band axis")
            return lo, max(lo, hi), False
        if isinstance(k, (int, np.integer)):
            i = int(k) + (rows if int(k) < 0 else 0)
            if not 0 <= i < rows:
                raise IndexError(f"BandedPinned: index {int(k)} outside [0, {rows})")
            return i, i + 1, True
        raise NotImplementedError(f"BandedPinned: key of type {type(k).__name__} on the band axis")

    def _pieces(self, lo, hi):
        """(band, first row in band, end row in band, offset of that first row from lo) covering [lo, hi)."""
        r = lo
        while r < hi:
            b, off = divmod(r, self.band_rows)
            end = min(hi - r, self.band_rows - off)
            yield b, off, off + end, r - lo
            r += end

    def _band_key(self, key, b0, b1, scalar):
        k = list(key)
        k[self.band_axis] = b0 if scalar else slice(b0, b1)
        return tuple(k)

    def _along(self, a0, a1):
        k = [slice(None)] * self.ndim
        k[self.band_axis] = slice(a0, a1)
        return tuple(k)

    def __getitem__(self, key):
        key = self._key(key)
        lo, hi, scalar = self._span(key[self.band_axis])
        parts = [self.bands[b][self._band_key(key, b0, b1, scalar)] for b, b0, b1, _ in self._pieces(lo, hi)]
        if len(parts) == 1:
            return parts[0]
        return np.concatenate(parts, axis=self.band_axis - sum(isinstance(k, (int, np.integer))
                                                                 for k in key[:self.band_axis]))

    def __setitem__(self, key, value):
        key = self._key(key)
        lo, hi, scalar = self._span(key[self.band_axis])
        if np.isscalar(value):
            for b, b0, b1, _ in self._pieces(lo, hi):
                self.bands[b][self._band_key(key, b0, b1, scalar)] = value
            return
        value = np.asarray(value)
        axis = self.band_axis - sum(isinstance(k, (int, np.integer)) for k in key[:self.band_axis])
        for b, b0, b1, o in self._pieces(lo, hi):
            src = value if scalar else value[(slice(None),) * axis + (slice(o, o + b1 - b0),)]
            self.bands[b][self._band_key(key, b0, b1, scalar)] = src

    def copy_to(self, dst, key=()):
        """dst[...] = self[key], band by band (key slices every axis; dst has the slice's shape)."""
        key = self._key(key)
        lo, hi, _ = self._span(key[self.band_axis])
        for b, b0, b1, o in self._pieces(lo, hi):
            dst[self._along(o, o + b1 - b0)] = self.bands[b][self._band_key(key, b0, b1, False)]

    def copy_from(self, src, key=()):
        """self[key] = src, band by band."""
        key = self._key(key)
        lo, hi, _ = self._span(key[self.band_axis])
        for b, b0, b1, o in self._pieces(lo, hi):
            self.bands[b][self._band_key(key, b0, b1, False)] = src[self._along(o, o + b1 - b0)]

    def copy_from_gpu(self, src, key=()):
        """self[key] = device tensor src: each band's piece is downloaded into its page-locked block."""
        key = self._key(key)
        lo, hi, _ = self._span(key[self.band_axis])
        for b, b0, b1, o in self._pieces(lo, hi):
            self._blocks[b][self._band_key(key, b0, b1, False)].copy_(src[self._along(o, o + b1 - b0)])

    def copy_to_gpu(self, dst, key=()):
        """Device tensor dst[...] = self[key]: each band's piece is uploaded from its page-locked block."""
        key = self._key(key)
        lo, hi, _ = self._span(key[self.band_axis])
        for b, b0, b1, o in self._pieces(lo, hi):
            piece = self._blocks[b][self._band_key(key, b0, b1, False)]
            dst[self._along(o, o + b1 - b0)].copy_(piece, non_blocking=False)


class StreamPipe:
    """Pipelined batch loop over fixed-size device slots.

    ``in_shape`` / ``out_shape`` are the largest batch's shapes; a batch may use
    a leading sub-range (``load`` returns the filled view, ``compute`` gets
    matching views)."""

    def __init__(self, device, in_shape, in_dtype, out_shape, out_dtype):
        self.device = torch.device(device)
        self.in_shape, self.out_shape = tuple(in_shape), tuple(out_shape)
        self.pin_in = [pinned_empty(self.in_shape, in_dtype) for _ in range(2)]
        self.pin_out = [pinned_empty(self.out_shape, out_dtype) for _ in range(2)]
        self.dev_in = [torch.empty(self.in_shape, dtype=in_dtype, device=self.device) for _ in range(2)]
        self.dev_out = [torch.empty(self.out_shape, dtype=out_dtype, device=self.device) for _ in range(2)]
        self.s_h2d, self.s_comp, self.s_d2h = (torch.cuda.Stream(self.device) for _ in range(3))
        self.ev_h2d = [torch.cuda.Event() for _ in range(2)]
        self.ev_comp = [torch.cuda.Event() for _ in range(2)]
        self.ev_d2h = [torch.cuda.Event() for _ in range(2)]
        torch.cuda.synchronize(self.device)  # slots exist before any side stream touches them

    def run(self, n_iter, load, compute, store):
        """load(k, pinned) -> view filled with batch k; compute(k, din, dout_full) -> device view of the result;
        store(k, pinned_view) copies batch k's result out of the pinned slot."""
        pending = None  # (k, slot, n_out_view_shape)
        for k in range(n_iter + 1):
            if k < n_iter:
                s = k % 2
                self.ev_h2d[s].synchronize()  # the upload of batch k - 2 has left this pinned slot
                hin = load(k, self.pin_in[s])
                din = self.dev_in[s].view(-1)[: hin.numel()].view(hin.shape)
                with torch.cuda.stream(self.s_h2d):
                    self.s_h2d.wait_event(self.ev_comp[s])  # batch k - 2 no longer reads this device slot
                    din.copy_(hin, non_blocking=True)
                    self.ev_h2d[s].record(self.s_h2d)
                with torch.cuda.stream(self.s_comp):
                    self.s_comp.wait_event(self.ev_h2d[s])
                    self.s_comp.wait_event(self.ev_d2h[s])  # batch k - 2's result has been downloaded
                    dout = compute(k, din, self.dev_out[s])
                    self.ev_comp[s].record(self.s_comp)
                hout = self.pin_out[s].view(-1)[: dout.numel()].view(dout.shape)
                with torch.cuda.stream(self.s_d2h):
                    self.s_d2h.wait_event(self.ev_comp[s])
                    hout.copy_(dout, non_blocking=True)
                    self.ev_d2h[s].record(self.s_d2h)
            if pending is not None:
                pk, ps, pview = pending
                self.ev_d2h[ps].synchronize()
                store(pk, pview)
            pending = (k, k % 2, hout) if k < n_iter else None
        torch.cuda.synchronize(self.device)

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

import torch

from model_mosaic_amd._carriers import pinned_empty


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

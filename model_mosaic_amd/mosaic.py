"""MosaicBatch: R, psi, Fresnel D and |.|^2 over an uneven mosaic of tiles as one HIP graph.

The reference pipeline runs this chain per tile from its scripts: R per z
chunk (step2_radon.py:147,191), then psi = exp(-k beta_ratio p) exp(i k p),
D and |.|^2 per angle chunk (step3_propagation.py:175,223, :94-110).  A mosaic
scan has tiles of different detector widths, angle counts and z stacks
(README.md:28-34: 4 x-tiles x 9 z-stacks).  MosaicBatch enqueues every tile's
chain as one graph (mm_mosaic_* in libmosaic_hip.so): tiles of equal geometry
share plans; the tiles are split over ``branches`` concurrent graph branches,
each with its own workspaces and stream.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from model_mosaic_amd import _lib
from model_mosaic_amd._carriers import require_gpu


class MosaicBatch:
    """An uneven batch of tiles ``(n, nz, theta)``: delta (nz, n, n) -> intensity (ntheta, nz, n).

    ``k`` and ``beta_ratio`` follow step3_propagation.py:94-103 (``inv_norm``
    and ``1 / beta_r``): psi = exp(-k beta_ratio p) exp(i k p) for p = R delta.
    ``graph=False`` enqueues the launches directly; ``autotune=True`` picks
    each Radon plan's gather tile order (and, for 4096-point columns, the
    column split) on its first tile at the first run; ``branches`` (1 - 8,
    default 2) is the number of concurrent graph branches the tiles are
    balanced over (results do not depend on it).
    """

    def __init__(self, tiles, wavelength, voxelsize, distance, k, beta_ratio, pad: str = "symmetric",
                 graph: bool = True, autotune: bool = False, branches: int = 2):
        require_gpu()
        self._lib = _lib.load()
        if pad not in ("symmetric", "zero"):
            raise ValueError("MosaicBatch: pad must be 'symmetric' or 'zero'")
        self.tiles = [(int(n), int(nz), np.asarray(th, dtype=np.float32)) for n, nz, th in tiles]
        if not self.tiles:
            raise ValueError("MosaicBatch: no tiles")
        nt = len(self.tiles)
        ns = (C.c_int * nt)(*[t[0] for t in self.tiles])
        nzs = (C.c_int * nt)(*[t[1] for t in self.tiles])
        nths = (C.c_int * nt)(*[len(t[2]) for t in self.tiles])
        th = np.ascontiguousarray(np.concatenate([t[2] for t in self.tiles]).astype(np.float64))
        if not 1 <= int(branches) <= 8:
            raise ValueError("MosaicBatch: branches must be in [1, 8]")
        flags = ((0 if graph else _lib.MM_MOSAIC_NO_GRAPH) | (_lib.MM_MOSAIC_AUTOTUNE if autotune else 0)
                 | _lib.MM_MOSAIC_BRANCHES(int(branches)))
        plan = C.c_void_p()
        _lib.check(self._lib.mm_mosaic_create(C.byref(plan), nt, ns, nzs, nths, th.ctypes.data_as(C.POINTER(C.c_double)),
                                              float(wavelength), float(voxelsize), float(distance), float(k),
                                              float(beta_ratio), 0 if pad == "symmetric" else 1, flags), "MosaicBatch")
        self._plan = plan
        self.device = torch.device("cuda", torch.cuda.current_device())

    def info(self) -> dict:
        f = self._lib.mm_mosaic_info
        return {"radon_plans": f(self._plan, _lib.MM_MOSAIC_INFO_RADON_PLANS),
                "fresnel_plans": f(self._plan, _lib.MM_MOSAIC_INFO_FRESNEL_PLANS),
                "graph_captures": f(self._plan, _lib.MM_MOSAIC_INFO_CAPTURES),
                "tiles": f(self._plan, _lib.MM_MOSAIC_INFO_TILES),
                "branches": f(self._plan, _lib.MM_MOSAIC_INFO_BRANCHES),
                "tile_branch": [self._lib.mm_mosaic_tile_knob(self._plan, t, _lib.MM_MOSAIC_TILE_BRANCH)
                                for t in range(len(self.tiles))],
                "gather_xcd_group": [self._lib.mm_mosaic_tile_knob(self._plan, t, _lib.MM_RADON_TUNE_GATHER_XCD_GROUP)
                                     for t in range(len(self.tiles))],
                "col_split": [self._lib.mm_mosaic_tile_knob(self._plan, t, _lib.MM_RADON_TUNE_COL_SPLIT)
                              for t in range(len(self.tiles))],
                "gather_root": [self._lib.mm_mosaic_tile_knob(self._plan, t, _lib.MM_RADON_TUNE_GATHER_ROOT)
                                for t in range(len(self.tiles))],
                "gather_zp": [self._lib.mm_mosaic_tile_knob(self._plan, t, _lib.MM_RADON_TUNE_GATHER_ZP)
                              for t in range(len(self.tiles))],
                "workspace_bytes": int(self._lib.mm_mosaic_workspace_bytes(self._plan))}

    def out_buffers(self):
        return [torch.empty((len(th), nz, n), dtype=torch.float32, device=self.device) for n, nz, th in self.tiles]

    def run(self, deltas, outs=None):
        """Intensities |D psi(R delta_t)|^2 of every tile (list of (ntheta, nz, n) float32 tensors)."""
        if not self._plan:
            raise RuntimeError("MosaicBatch: plan was freed")
        if len(deltas) != len(self.tiles):
            raise ValueError(f"MosaicBatch.run: {len(deltas)} inputs for {len(self.tiles)} tiles")
        outs = self.out_buffers() if outs is None else outs
        for t, ((n, nz, th), d, o) in enumerate(zip(self.tiles, deltas, outs)):
            if d.dtype != torch.float32 or tuple(d.shape) != (nz, n, n) or not d.is_contiguous() or d.device != self.device:
                raise ValueError(f"MosaicBatch.run: tile {t} input must be contiguous float32 ({nz}, {n}, {n}) on {self.device}")
            if (o.dtype != torch.float32 or tuple(o.shape) != (len(th), nz, n) or not o.is_contiguous()
                    or o.device != self.device):
                raise ValueError(f"MosaicBatch.run: tile {t} output must be contiguous float32 ({len(th)}, {nz}, {n}) "
                                 f"on {self.device}")
        nt = len(self.tiles)
        din = (C.c_void_p * nt)(*[d.data_ptr() for d in deltas])
        dout = (C.c_void_p * nt)(*[o.data_ptr() for o in outs])
        _lib.check(self._lib.mm_mosaic_run(self._plan, din, dout, _lib.stream_handle()), "MosaicBatch.run")
        return outs

    def free(self):
        if getattr(self, "_plan", None):
            torch.cuda.synchronize(self.device)
            self._lib.mm_mosaic_destroy(self._plan)
            self._plan = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

"""MI355X-native (gfx950) USFFT Radon transform and Fresnel propagation.

Drop-in replacements for model_mosaic's hot-path operators:

* ``model_mosaic_amd.processing.tomo.TomoReal``          (processing/tomo.py:17)
* ``model_mosaic_amd.processing.tomo_large.TomoLargeReal`` (processing/tomo_large.py:46)
* ``model_mosaic_amd.processing.propagation.Propagation``  (processing/propagation.py:15), plus ``DT``
* ``model_mosaic_amd.processing.propagation_large.PropagationLarge`` (processing/propagation_large.py:35)
* ``model_mosaic_amd.step_kernels.psi_from_proj`` / ``abs2`` (step3_propagation.py:94-110)
* ``model_mosaic_amd.model.ForwardModel``: differentiable (delta, beta) -> |D psi|^2 misfit.

All compute runs in libmosaic_hip.so (hand-written HIP for gfx950) through the
C ABI of include/mosaic_hip.h; torch tensors are only the carriers.
"""
__version__ = "0.1.0"

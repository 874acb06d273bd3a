"""CPU oracle for the USFFT Radon transform and Fresnel propagation.

TEST INFRASTRUCTURE ONLY.  This package checks results; it is never the product:
only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it.  The product path (``model_mosaic_amd``) never
imports, links or executes anything from here and fails loudly when its HIP
library is missing.

Modules
-------
usfft        NumPy restatement of ``processing/tomo.py`` + ``processing/kernels.py``
             (``TomoReal.R`` / ``TomoReal.RT``) and of the arithmetic of
             ``processing/tomo_large.py`` (``TomoLargeReal``), in FP32
             ("reference restatement") and FP64 ("exact arithmetic") modes.
fresnel      NumPy restatement of ``processing/propagation.py`` (``Propagation.D``),
             ``processing/propagation_large.py`` (separable ``PropagationLarge.D``),
             the derived adjoint ``DT`` and the ``step3_propagation.py`` psi / |.|^2
             element-wise kernels.
independent  Oracles that do not share code with the restatement: exact polar
             NUDFT in FP64, analytic projections of Gaussian blobs, the
             cosine-grating Fresnel closed form.
model        FP64 forward model L(delta, beta) and its gradient, built from the
             restatement, plus a central finite-difference checker.

How each function is pinned (the "parity pinned" evidence) is listed in each
module's header.  No reference golden vectors exist (the reference's tests are
relational only), so pins are closed forms, brute force and identities.
"""

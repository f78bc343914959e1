"""Learned warm start for a parametric LP family: GPU PDLP solves as training
data, a PCA basis of the solutions, and a BF16 MFMA MLP (hand-written gfx950
kernels) from the family parameter to PCA coefficients.  Cold-start PDLP stays
the fallback and the judge of accuracy."""
from .family import BlockFamily

__all__ = ["BlockFamily"]

"""MI355X-native permanent engine: a drop-in for piquasso's
``permanent(matrix, rows, cols)`` (piquasso/_math/permanent.cpp:71-78) whose
Glynn / Gray-code sweep runs in hand-written HIP kernels for gfx950.

Public entry points
-------------------
permanent(matrix, rows, cols)        drop-in for piquasso._math.permanent.permanent
permanent_ex(matrix, rows, cols, mode) one permanent with the double-double parts
permanent_exact(matrix, rows, cols)  the exact integer permanent of an integer matrix
PermanentBatch / permanent_batch     many uneven permanents in one persistent launch
permanent_laplace(matrix, rows, cols) drop-in for piquasso._math.permanent.permanent_laplace
permanent_grad / permanent_grad_batch d perm / d A (grad_perm, any shape)
torch_permanent(A, rows, cols)       differentiable permanent (torch.autograd.Function)
torch_permanent_batch(A, rows, cols) the same for a (B, m_r, m_c) stack, one engine call
output_probabilities(U, s, outputs)  boson-sampling probabilities of many outputs in one batch
particle_number_probabilities(...)   the same for a superposition of input Fock states
permanent_laplace_batch(ms, rs, cs)  many permanent_laplace problems in one launch
generate_samples / generate_lossy_samples  Clifford & Clifford sampling, shots in lockstep
tensor_permanent / detection_probabilities  uniformly partially distinguishable photons
lossy_detection_probabilities        any Gram matrix, lossy T (matrix-polynomial permanent)
generate_lossy_and_partially_distinguishable_samples  sampling from those probabilities
marginal_fock_probabilities / generate_marginal_samples  marginals from permanent-polynomial moments
"""

from ._math.permanent import (permanent, permanent_exact, permanent_laplace,
                              permanent_laplace_batch)
from .engine import PermanentBatch, permanent_batch, permanent_ex
from .grad import (PermanentBatchFunction, PermanentFunction, permanent_grad, permanent_grad_batch,
                   torch_permanent, torch_permanent_batch)
from .probabilities import (output_amplitudes, output_probabilities,
                            particle_number_probabilities)
from .distinguishability import (detection_probabilities, detection_probability, general_input_norm,
                                 lossy_detection_probabilities, perm_poly_coefficients,
                                 tensor_permanent, tensor_permanent_batch)
from .sampling import (fock_space_basis, generate_lossy_and_partially_distinguishable_samples,
                       generate_lossy_samples, generate_samples, postselected_fock_basis)
from .marginal import generate_marginal_samples, marginal_fock_probabilities

__all__ = [
    "permanent", "permanent_exact", "permanent_laplace", "permanent_ex", "PermanentBatch", "permanent_batch",
    "permanent_grad", "permanent_grad_batch", "PermanentFunction", "torch_permanent",
    "output_amplitudes", "output_probabilities", "permanent_laplace_batch",
    "generate_samples", "generate_lossy_samples", "tensor_permanent", "tensor_permanent_batch",
    "detection_probability", "detection_probabilities", "torch_permanent_batch",
    "PermanentBatchFunction", "perm_poly_coefficients", "general_input_norm",
    "lossy_detection_probabilities", "generate_lossy_and_partially_distinguishable_samples",
    "fock_space_basis", "postselected_fock_basis", "particle_number_probabilities",
    "marginal_fock_probabilities", "generate_marginal_samples",
]
__version__ = "0.1.0"

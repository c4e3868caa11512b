"""Learned mesh-CNN potential correction on the HIP library (PM++ ``pmpp.corrections``)."""

from .conv import conv3d, conv3d_bwd_data, conv3d_bwd_weight  # noqa: F401
from .mesh_cnn import (MeshCNNPotentialCorrection, evaluate_residual, init_mesh_cnn_potential_correction,  # noqa: F401
                       param_layout, residual_backward)
from .training import (build_correction_optimizer, corrected_gravity, density, force_loss,  # noqa: F401
                       gravity_with_residual, gravity_with_residual_vjp, train_step)

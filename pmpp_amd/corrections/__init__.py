"""Learned mesh-CNN potential correction on the HIP library (PM++ ``pmpp.corrections``)."""

from .conv import conv3d, conv3d_bwd_data, conv3d_bwd_weight  # noqa: F401
from .mesh_cnn import (MeshCNNPotentialCorrection, correction_meta, evaluate_residual,  # noqa: F401
                       init_mesh_cnn_potential_correction, load_correction, param_layout, residual_backward,
                       save_correction, with_potential_channel)
from .training import (build_correction_optimizer, compare_runs, corrected_gravity, density,  # noqa: F401
                       fine_run_pairs, force_loss, gravity_with_residual, gravity_with_residual_vjp, mean_force_loss,
                       pk_errors, train_correction, train_step)

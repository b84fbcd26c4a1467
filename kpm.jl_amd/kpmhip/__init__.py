"""kpmhip -- MI355X (gfx950) Chebyshev moment engine behind KPM.jl's moment API.

Compute runs in libkpm_hip.so (hand-written HIP kernels, C ABI in
include/kpm_hip.h); this package is the host-side mirror of the reference's
interface: the moment engine (src/moment.jl), its primitives
(src/utils/chebyshev_iteration.jl, src/utils/vectors.jl), the device layer
(src/device.jl), the rescaling (src/utils/external.jl), damping kernels
(src/kernels), the DOS / Kubo–Bastin reconstruction heads, the typed front
end (src/frontend.jl), and the matrix-function action with its consumers
(src/utils/chebyshev_action.jl: evolve, fermi_projector, chern_marker).
"""
from ._lib import (ArgumentError, DomainError, KPMHipError, KPMHipMissing, KPMUnstableError,
                   QuadratureError, load_library)
from .action import (chebyshev_action, chebyshev_action_, chern_marker, chern_marker_average,
                     chern_marker_region, evolution_coefficients, evolution_order,
                     evolution_tail, evolve, evolve_, fermi_coefficients, fermi_projector)
from .adjoint import (AdjointTape, bond_gradient, dos_loss, hopping_gradient, kpm_1d_adjoint,
                      kpm_1d_adjoint_forward, kpm_2d_adjoint, kubo_bastin_weights, moments_rrule,
                      sigma_rrule, site_gradient)
from .applications import (Gamma_nm, Gamma_nm_mu, chebyshev_lin_trans, chebyshevT,
                           chebyshevT_0, chebyshevT_accurate, chebyshevT_xn, d_dc_cond, dc_cond0,
                           dc_cond_single, fermiFunction, fermiFunctions, gausschebyshevt)
from .applications import dos as dos_raw
from .applications import dos0 as dos0_raw
from .applications import kubo_bastin_cond as kubo_bastin_cond_raw
from .batch import BatchOperator, kpm_batch_1d
from .bdg import (BdGOperator, ScaledOperator, bdg_assemble, chebyshev_stability_probe,
                  gershgorin_bound, spectral_radius)
from .stiffness import (diamagnetic_term, nambu_current_q, nambu_diamagnetic, superfluid_stiffness,
                        two_energy_response)
from .cpge import Table3, cpge, d_cpge
from . import ensemble
from .device import (ACTIVE_DEVICE, CPUDevice, DeviceAccumulator, DeviceBlock, DeviceCSR, HIPDevice, activate,
                     device_free_memory, get_context, maybe_on_device_zeros, maybe_to_host,
                     to_device, whichcore)
from .frontend import (ConductivityMoments, CurrentMoments, DosMoments, GreenMoments,
                       RescaledHamiltonian, cond_moments, current_moments, dos, dos0, dos_moments,
                       green_moments, kubo_bastin_cond, ldos_moments, nc, npairs,
                       random_phase_vectors, rescale)
from .kernels import (JacksonKernel, LorentzKernels, hn, mu2D_apply_kernel_and_h,
                      mu3D_apply_kernel_and_h, muND_apply_kernel_and_h)
from .moment import (engine_blocking, kpm_1d, kpm_1d_, kpm_1d_current, kpm_1d_current_,
                     kpm_2d, kpm_2d_, kpm_2d_blocking, kpm_3d, kpm_3d_)
from .primitives import (chebyshev_iter, chebyshev_iter_single, chebyshev_iter_wrap,
                         columnwise_dot, copy_moment_accumulator, moment_accumulator,
                         normalize_by_col, r2_i, r2_ip,
                         r2_ipp, r_i, r_ip, r_ipp, spmm)
from .rescale import normalizeH
from .rng import Xoshiro256pp, default_rng
from .tune import autotune, tune_get, tune_set, tune_set_chunked

__all__ = [n for n in dir() if not n.startswith("_")]

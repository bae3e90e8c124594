"""cfft_hip — host side of the MI355X KRFFT / pseudo-spectral diffusion library.

Mirrors the plan/execute API of liuyxpp/CrystallographicFFT.jl for the hot path
and calls the C ABI of ``libcfft_hip.so`` (include/cfft_hip.h).  Julia's
mutating ``f!`` functions are exposed as ``f`` and mutate their arguments.
"""
from ._lib import (CFFTError, CFFTLibraryMissing, FP64, MIXED, PATH_AUTO, PATH_M2, PATH_M2Q,
                   PATH_M7, PATH_M7Q, PATH_QR, PATH_QS, lib)
from .symmetry_ops import (CentA, CentC, CentF, CentI, CentP, CenteringType, SymOp, apply_op,
                           check_shift_invariance, convert_op, detect_centering_type, dual_ops,
                           get_ops, ops_arrays, pack_ops, strip_centering)
from .plans import (ASUPoint, CenteredKRFFTBackwardPlan, CenteredKRFFTPlan, CenteredSCFTPlan,
                    CenteringPreFoldKRFFTPlan,
                    DevicePlan, GeneralCFFTPlan, M2BackwardPlan, M2QPlan, M2SCFTPlan,
                    M7SCFTPlan, SpectralIndexing, auto_L, build_fill_map, calc_spectral_asu,
                    count_special_positions, detect_centering, execute_centered_ikrfft,
                    execute_centered_krfft, execute_centered_scft, execute_krfft,
                    execute_m2_backward, execute_m2_q, execute_m2_scft, execute_m7_scft,
                    fast_reconstruct, fft_reconstruct, fft_reconstruct_centered,
                    find_optimal_shift, fullgrid_to_subgrid, get_k_vector,
                    ifft_unrecon_centered, pack_stride, pack_stride_real, plan_centered_ikrfft,
                    plan_centered_scft, plan_krfft, plan_krfft_centered, plan_m2_backward,
                    plan_m2_q, plan_m2_scft, plan_m7_scft, select_rep_ops, subgrid_to_fullgrid,
                    update_kernel, update_m2_kernel)
from .asu import (ASUBlock, CFFTPlan, CrystallographicASU, calc_asu, ldiv, mul, pack_asu,
                  plan_cfft)
from . import scft  # noqa: F401
from .diffusion_solver import (DiffusionSolver, QFusedDiffusionSolver, apply_diffusion_operator,
                               interaction_step, plan_diffusion, step_diffusion)

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
from .asu import (ASUBlock, CFFTPlan, CrystallographicASU, InterleavedOrbit, analyze_interleaved_orbits,
                  calc_asu, ldiv, mul, pack_asu, pack_asu_interleaved, plan_cfft)
from . import scft  # noqa: F401
from .families import (FractalCFFTPlan, G0ASUBackwardPlan, G0ASUPlan, GeneralG0ASUPlan,
                       RecursiveKRFFTPlan, SelectiveG0Plan, SparseKRFFTPlan, StagedPm3mPlan,
                       compute_spectral_asu, execute_fractal_krfft, execute_fractal_krfft_v2,
                       execute_fractal_krfft_v3, execute_g0asu_ikrfft, execute_g0asu_krfft,
                       execute_general_g0asu_krfft, execute_recursive_krfft,
                       execute_selective_krfft, execute_sparse_krfft, execute_staged,
                       execute_staged_full, has_cubic_p3c_symmetry, lookup_full_spectrum, shift_ops_half_grid,
                       plan_fractal_krfft, plan_fractal_krfft_v2, plan_fractal_krfft_v3,
                       plan_krfft_g0asu, plan_krfft_g0asu_backward, plan_krfft_g0asu_general,
                       plan_krfft_recursive, plan_krfft_selective, plan_krfft_sparse,
                       plan_staged_pm3m, plan_staged_pm3m_full)
from .diffusion_solver import (DiffusionSolver, MatrixDiffusionSolver, QFusedDiffusionSolver,
                               apply_diffusion_operator, calc_transform_matrix, interaction_step,
                               plan_diffusion, step_diffusion)

# Julia binding of libcfft_hip.so for CrystallographicFFT.jl users.
#
# A maintainer adds this module beside the reference package.  It keeps the
# reference's plan/execute names, argument order and meaning, return values
# and error messages, and forwards the hot path to the C ABI declared in
# include/cfft_hip.h (host-pointer entry points: blocking, FP64).  Symmetry
# set-up (get_ops, find_optimal_shift, calc_spectral_asu, calc_asu, pack_asu)
# stays the reference's own Julia code.  Julia is not installed in the build
# image, so this file is written to the ABI but has not been run.
#
# Reference call -> GPU plan path
#   plan_m2_q / execute_m2_q!                 -> CFFT_PATH_M2Q   cfft_diffuse
#   plan_m7_scft / execute_m7_scft!           -> CFFT_PATH_M7Q   cfft_diffuse
#   plan_m2_scft / execute_m2_scft!           -> CFFT_PATH_M2    cfft_diffuse
#   plan_centered_scft / execute_centered_scft! -> CFFT_PATH_M7  cfft_diffuse
#   update_m2_kernel! / update_kernel!        -> cfft_update_kernel
#   plan_krfft(spec, ops) / fft_reconstruct!  -> CFFT_PATH_M2    cfft_forward
#   plan_m2_backward / execute_m2_backward!   -> CFFT_PATH_M2    cfft_backward
#   plan_krfft_centered / execute_centered_krfft! / fft_reconstruct_centered!
#                                             -> CFFT_PATH_M7    cfft_forward
#   plan_centered_ikrfft / execute_centered_ikrfft! -> CFFT_PATH_M7 cfft_backward
#   fullgrid_to_subgrid! / subgrid_to_fullgrid! -> cfft_full_to_sub / cfft_sub_to_full
#   plan_diffusion(...; method=:qfused|:m2|:centered|:qs) / step_diffusion! /
#   apply_diffusion_operator!                 -> cfft_step / cfft_diffuse
#   plan_diffusion(...; method=:matrix|:krfft) -> the reference's own CPU solvers
#   plan_cfft / mul! / ldiv!                  -> cfft_block_plan_create / cfft_block_forward / cfft_block_inverse
module CrystallographicFFTHIP

using LinearAlgebra
import LinearAlgebra: mul!, ldiv!
import CrystallographicFFT   # qualified use only: the names below are this module's own
using CrystallographicFFT.SymmetryOps: SymOp, get_ops, detect_centering_type, CentP
using CrystallographicFFT.ASU: find_optimal_shift, calc_asu, pack_asu, CrystallographicASU
using CrystallographicFFT.SpectralIndexing: SpectralIndexing, calc_spectral_asu
import CrystallographicFFT.DiffusionSolver

export plan_m2_q, execute_m2_q!, plan_m7_scft, execute_m7_scft!
export plan_m2_scft, execute_m2_scft!, update_m2_kernel!
export plan_centered_scft, execute_centered_scft!, update_kernel!
export plan_krfft, fft_reconstruct!, pack_stride!, plan_m2_backward, execute_m2_backward!
export plan_krfft_centered, execute_centered_krfft!, fft_reconstruct_centered!
export plan_centered_ikrfft, execute_centered_ikrfft!
export fullgrid_to_subgrid!, subgrid_to_fullgrid!
export plan_diffusion, step_diffusion!, apply_diffusion_operator!
export plan_cfft, CFFTError

const libcfft = get(ENV, "CFFT_HIP_LIB", joinpath(@__DIR__, "..", "libcfft_hip.so"))

const PATH_AUTO, PATH_M2, PATH_M7, PATH_M2Q, PATH_M7Q, PATH_QR, PATH_QS = Int32.(0:6)
const FP64, MIXED = Int32(0), Int32(1)
# plan options (cfft_plan_set_option); no counterpart in the reference
const OPT_PERSISTENT, OPT_FUSE, OPT_GRAPH, OPT_TILE_ROWS, OPT_FP32_TWIDDLES, OPT_FFT_LANES,
      OPT_PAIR_THREADS, OPT_AXIS_SPLIT = Int32.(1:8)

struct CFFTError <: Exception
    code::Int32
    msg::String
end
Base.showerror(io::IO, e::CFFTError) = print(io, "cfft_hip error ", e.code, ": ", e.msg)

function check(status::Cint)
    status == 0 && return nothing
    buf = Vector{UInt8}(undef, 4096)
    ccall((:cfft_last_error, libcfft), Cint, (Ptr{UInt8}, Csize_t), buf, length(buf))
    throw(CFFTError(status, unsafe_string(pointer(buf))))
end

# ops -> nops x 12 Int32 (R row-major, then t), the ABI layout
function pack_ops(ops::Vector{<:SymOp})
    out = Vector{Int32}(undef, 12 * length(ops))
    for (g, op) in enumerate(ops)
        b = 12 * (g - 1)
        for i in 1:3, j in 1:3
            out[b + 3 * (i - 1) + j] = Int32(op.R[i, j])
        end
        for i in 1:3
            out[b + 9 + i] = Int32(op.t[i])
        end
    end
    return out
end

# ---------------------------------------------------------------------------
# Device plan (one C plan: tables and work buffers on the GPU)
# ---------------------------------------------------------------------------

mutable struct DevicePlan
    handle::Ptr{Cvoid}
    N::NTuple{3,Int}
    M::NTuple{3,Int}
    L::NTuple{3,Int}
    n_spec::Int
    batch::Int
end

function DevicePlan(N::NTuple{3,Int}, ops_shifted, ds::Float64, lattice::AbstractMatrix;
                    path=PATH_AUTO, precision=FP64, batch=1)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    lat = Float64.(vec(Matrix(lattice)))            # column-major, columns a, b, c
    check(ccall((:cfft_plan_create, libcfft), Cint,
                (Ptr{Int32}, Ptr{Int32}, Int32, Float64, Ptr{Float64}, Int32, Int32, Int32,
                 Ptr{Ptr{Cvoid}}),
                Int32[N...], pack_ops(ops_shifted), Int32(length(ops_shifted)), ds, lat,
                Int32(path), Int32(precision), Int32(batch), h))
    info = zeros(Int32, 18)
    check(ccall((:cfft_plan_info, libcfft), Cint, (Ptr{Cvoid}, Ptr{Int32}), h[], info))
    p = DevicePlan(h[], N, (info[1], info[2], info[3]), (info[4], info[5], info[6]),
                   Int(info[7]), Int(info[13]))
    finalizer(p) do x
        x.handle == C_NULL || ccall((:cfft_plan_destroy, libcfft), Cint, (Ptr{Cvoid},), x.handle)
        x.handle = C_NULL
    end
    return p
end

# Launch schedule of a plan (e.g. set_option!(dev, OPT_AXIS_SPLIT, 1): the per-axis chain
# of path QS); results do not depend on it beyond rounding.
set_option!(dev::DevicePlan, key::Integer, value::Integer) =
    check(ccall((:cfft_plan_set_option, libcfft), Cint, (Ptr{Cvoid}, Int32, Int64),
                dev.handle, Int32(key), Int64(value)))

_ops_for(N, sg_num, dim) = find_optimal_shift(get_ops(sg_num, dim, N), N)[2]

function _check_f0(dev::DevicePlan, f0::AbstractArray)
    # src/q_fused_krfft.jl:998-1001: the only execute-time check is the f0 size
    size(f0) == dev.M || length(f0) == prod(dev.M) ||
        error("f0 size $(size(f0)) does not match plan subgrid dims $(dev.M).")
end

function _diffuse!(dev::DevicePlan, f0::Array{Float64})
    _check_f0(dev, f0)
    check(ccall((:cfft_diffuse, libcfft), Cint, (Ptr{Cvoid}, Ptr{Float64}), dev.handle, f0))
    return f0
end

function _forward!(dev::DevicePlan, f0::Array{Float64}, F::Vector{ComplexF64})
    _check_f0(dev, f0)
    length(F) == dev.n_spec || error("F_spec has $(length(F)) entries, spectral ASU has $(dev.n_spec)")
    check(ccall((:cfft_forward, libcfft), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{ComplexF64}),
                dev.handle, f0, F))
    return F
end

function _backward!(dev::DevicePlan, F::AbstractVector{ComplexF64}, f0::Array{Float64})
    _check_f0(dev, f0)
    length(F) == dev.n_spec || error("F_spec has $(length(F)) entries, spectral ASU has $(dev.n_spec)")
    Fc = F isa Vector{ComplexF64} ? F : Vector{ComplexF64}(F)
    check(ccall((:cfft_backward, libcfft), Cint, (Ptr{Cvoid}, Ptr{ComplexF64}, Ptr{Float64}),
                dev.handle, Fc, f0))
    return f0
end

function _update!(dev::DevicePlan, Δs::Float64, lattice::AbstractMatrix)
    check(ccall((:cfft_update_kernel, libcfft), Cint, (Ptr{Cvoid}, Float64, Ptr{Float64}),
                dev.handle, Δs, Float64.(vec(Matrix(lattice)))))
end

# ---------------------------------------------------------------------------
# Q-fused subgrid diffusion: src/q_fused_krfft.jl
# ---------------------------------------------------------------------------

"M2QPlan on the GPU: the fields the reference's callers read (L, M, N)."
struct M2QPlanHIP
    dev::DevicePlan
    L::NTuple{3,Int}
    M::NTuple{3,Int}
    N::NTuple{3,Int}
end

"plan_m2_q(N, sg_num, dim, Δs, lattice) (src/q_fused_krfft.jl:83-169)."
function plan_m2_q(N::Tuple, sg_num::Int, dim::Int, Δs::Float64, lattice::AbstractMatrix)
    @assert length(N) == dim == 3
    dev = DevicePlan(NTuple{3,Int}(N), _ops_for(N, sg_num, dim), Δs, lattice; path=PATH_M2Q)
    return M2QPlanHIP(dev, dev.L, dev.M, dev.N)
end

"execute_m2_q!(plan, f0) (src/q_fused_krfft.jl:553-576): f0 <- D f0 in place."
execute_m2_q!(plan::M2QPlanHIP, f0::Array{Float64}) = _diffuse!(plan.dev, f0)

"subgrid_to_fullgrid!(f_full, f0, plan) (src/q_fused_krfft.jl:681-688)."
function subgrid_to_fullgrid!(f_full::Array{Float64}, f0::Array{Float64}, plan::M2QPlanHIP)
    _check_f0(plan.dev, f0)
    length(f_full) == prod(plan.N) || error("f_full size $(size(f_full)) does not match N $(plan.N).")
    check(ccall((:cfft_sub_to_full, libcfft), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}),
                plan.dev.handle, f0, f_full))
    return f_full
end

"fullgrid_to_subgrid!(f0, f_full, plan) (src/q_fused_krfft.jl:698-719)."
function fullgrid_to_subgrid!(f0::Array{Float64}, f_full::Array{Float64}, plan::M2QPlanHIP)
    _check_f0(plan.dev, f0)
    length(f_full) == prod(plan.N) || error("f_full size $(size(f_full)) does not match N $(plan.N).")
    check(ccall((:cfft_full_to_sub, libcfft), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}),
                plan.dev.handle, f_full, f0))
    return f0
end

"M7SCFTPlan on the GPU (centring fold + Q multiply + unfold)."
struct M7SCFTPlanHIP
    dev::DevicePlan
    M::NTuple{3,Int}
end

"plan_m7_scft(N, sg_num, dim, Δs, lattice) (src/q_fused_krfft.jl:822-868)."
function plan_m7_scft(N::NTuple{3,Int}, sg_num::Int, dim::Int, Δs::Float64, lattice::AbstractMatrix)
    dev = DevicePlan(N, _ops_for(N, sg_num, dim), Δs, lattice; path=PATH_M7Q)
    return M7SCFTPlanHIP(dev, dev.M)
end

"execute_m7_scft!(plan, f0) (src/q_fused_krfft.jl:870-895)."
execute_m7_scft!(plan::M7SCFTPlanHIP, f0::Array{Float64,3}) = _diffuse!(plan.dev, f0)

"M2SCFTPlan on the GPU (forward KRFFT, spectral K, backward)."
struct M2SCFTPlanHIP
    dev::DevicePlan
    M::NTuple{3,Int}
end

"plan_m2_scft(N, sg_num, dim, Δs, lattice) (src/q_fused_krfft.jl:948-986)."
function plan_m2_scft(N::Tuple, sg_num::Int, dim::Int, Δs::Float64, lattice::AbstractMatrix)
    @assert length(N) == dim == 3
    dev = DevicePlan(NTuple{3,Int}(N), _ops_for(N, sg_num, dim), Δs, lattice; path=PATH_M2)
    return M2SCFTPlanHIP(dev, dev.M)
end

"execute_m2_scft!(plan, f0) (src/q_fused_krfft.jl:989-1026): f0 <- D f0 in place."
execute_m2_scft!(plan::M2SCFTPlanHIP, f0::Array{Float64,3}) = _diffuse!(plan.dev, f0)

"update_m2_kernel!(plan, N, sg_num, dim, Δs_new, lattice) (src/q_fused_krfft.jl:1036-1052)."
update_m2_kernel!(plan::M2SCFTPlanHIP, N::Tuple, sg_num::Int, dim::Int, Δs_new::Float64,
                  lattice::AbstractMatrix) = _update!(plan.dev, Δs_new, lattice)

# ---------------------------------------------------------------------------
# Centring-fold pipelines: src/centering_fold.jl
# ---------------------------------------------------------------------------

struct CenteredSCFTPlanHIP
    dev::DevicePlan
end

"plan_centered_scft(spec_asu, ops_shifted, N, Δs, lattice) (src/centering_fold.jl:868-894)."
plan_centered_scft(spec_asu::SpectralIndexing, ops_shifted::Vector{<:SymOp}, N::Tuple,
                   Δs::Float64, lattice::AbstractMatrix) =
    CenteredSCFTPlanHIP(DevicePlan(NTuple{3,Int}(N), ops_shifted, Δs, lattice; path=PATH_M7))

"execute_centered_scft!(plan, f0) (src/centering_fold.jl:908-931)."
execute_centered_scft!(plan::CenteredSCFTPlanHIP, f0::Array{Float64,3}) = _diffuse!(plan.dev, f0)

"update_kernel!(plan, spec_asu, N, Δs_new, lattice) (src/centering_fold.jl:939-953)."
update_kernel!(plan::CenteredSCFTPlanHIP, spec_asu::SpectralIndexing, N::Tuple, Δs_new::Float64,
               lattice::AbstractMatrix) = _update!(plan.dev, Δs_new, lattice)

"CenteredKRFFTPlan on the GPU: f0_buffer (real subgrid) and output_buffer as in the reference."
struct CenteredKRFFTPlanHIP
    dev::DevicePlan
    f0_buffer::Array{Float64,3}
    output_buffer::Vector{ComplexF64}
end

"plan_krfft_centered(spec_asu, ops_shifted; centering=:auto) (src/centering_fold.jl:328-389)."
function plan_krfft_centered(spec_asu::SpectralIndexing, ops_shifted::Vector{<:SymOp}; centering=:auto)
    N = NTuple{3,Int}(spec_asu.N)
    cent = centering === :auto ? detect_centering_type(ops_shifted, N) : centering
    cent == CentP && error("No centering detected. Use plan_krfft instead.")
    dev = DevicePlan(N, ops_shifted, 0.0, Matrix{Float64}(I, 3, 3); path=PATH_M7)
    return CenteredKRFFTPlanHIP(dev, zeros(Float64, dev.M), zeros(ComplexF64, dev.n_spec))
end

"execute_centered_krfft!(plan, u) (src/centering_fold.jl:392-413): full real grid -> spectral ASU."
function execute_centered_krfft!(plan::CenteredKRFFTPlanHIP, u::AbstractArray{<:Real,3})
    size(u) == plan.dev.N || error("u size $(size(u)) does not match N $(plan.dev.N).")
    uf = u isa Array{Float64,3} ? u : Array{Float64,3}(u)
    check(ccall((:cfft_full_to_sub, libcfft), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}),
                plan.dev.handle, uf, plan.f0_buffer))
    return _forward!(plan.dev, plan.f0_buffer, plan.output_buffer)
end

"fft_reconstruct_centered!(plan) (src/centering_fold.jl:421-431): reads plan.f0_buffer."
fft_reconstruct_centered!(plan::CenteredKRFFTPlanHIP) =
    _forward!(plan.dev, plan.f0_buffer, plan.output_buffer)

struct CenteredKRFFTBackwardPlanHIP
    dev::DevicePlan
end

"plan_centered_ikrfft(spec_asu, ops_shifted, fwd_plan) (src/centering_fold.jl:630-731)."
plan_centered_ikrfft(spec_asu::SpectralIndexing, ops_shifted::Vector{<:SymOp},
                     fwd_plan::CenteredKRFFTPlanHIP) = CenteredKRFFTBackwardPlanHIP(fwd_plan.dev)

"execute_centered_ikrfft!(bplan, F_spec, f0_out) (src/centering_fold.jl:745-770)."
execute_centered_ikrfft!(bplan::CenteredKRFFTBackwardPlanHIP, F_spec::AbstractVector{ComplexF64},
                         f0_out::Array{Float64,3}) = _backward!(bplan.dev, F_spec, f0_out)

# ---------------------------------------------------------------------------
# General KRFFT forward / M2 backward: src/krfft.jl
# ---------------------------------------------------------------------------

"GeneralCFFTPlan on the GPU: input_buffer (M^3, complex) -> output_buffer (spectral ASU)."
struct GeneralCFFTPlanHIP
    dev::DevicePlan
    input_buffer::Vector{ComplexF64}
    output_buffer::Vector{ComplexF64}
    grid_N::NTuple{3,Int}
    subgrid_dims::NTuple{3,Int}
    f0::Array{Float64,3}      # real part of input_buffer handed to the device
end

"plan_krfft(spec_asu, ops_shifted) (src/krfft.jl:301-389)."
function plan_krfft(spec_asu::SpectralIndexing, ops_shifted::Vector{<:SymOp})
    N = NTuple{3,Int}(spec_asu.N)
    dev = DevicePlan(N, ops_shifted, 0.0, Matrix{Float64}(I, 3, 3); path=PATH_M2)
    return GeneralCFFTPlanHIP(dev, zeros(ComplexF64, prod(dev.M)), zeros(ComplexF64, dev.n_spec),
                              N, dev.M, zeros(Float64, dev.M))
end

"""fft_reconstruct!(plan) (src/krfft.jl:554-562): FFT of plan.input_buffer and the
spectral-ASU reconstruction; returns plan.output_buffer (overwritten by the next
call, as in the reference).  The device path transforms real subgrid fields (the
symmetric fields every caller passes as complex(f0)); an input with a nonzero
imaginary part is refused."""
function fft_reconstruct!(plan::GeneralCFFTPlanHIP)
    any(x -> imag(x) != 0.0, plan.input_buffer) &&
        throw(ArgumentError("input_buffer must hold a real field (imaginary parts zero)"))
    plan.f0 .= reshape(real.(plan.input_buffer), plan.subgrid_dims)
    return _forward!(plan.dev, plan.f0, plan.output_buffer)
end

"pack_stride!(plan, u) (src/krfft.jl:571-600): u[1:L:end, ...] into plan.input_buffer."
function pack_stride!(plan::GeneralCFFTPlanHIP, u::AbstractArray{<:Real})
    uf = u isa Array{Float64,3} ? u : Array{Float64,3}(u)
    check(ccall((:cfft_full_to_sub, libcfft), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}),
                plan.dev.handle, uf, plan.f0))
    plan.input_buffer .= vec(complex.(plan.f0))
    return plan
end

"M2BackwardPlan on the GPU: f0_buf (complex subgrid) as in the reference."
struct M2BackwardPlanHIP
    dev::DevicePlan
    f0_buf::Array{ComplexF64,3}
    f0_real::Array{Float64,3}
end

"plan_m2_backward(spec_asu, ops_shifted) (src/krfft.jl:1101-1170)."
function plan_m2_backward(spec_asu::SpectralIndexing, ops_shifted::Vector{<:SymOp})
    N = NTuple{3,Int}(spec_asu.N)
    dev = DevicePlan(N, ops_shifted, 0.0, Matrix{Float64}(I, 3, 3); path=PATH_M2)
    return M2BackwardPlanHIP(dev, zeros(ComplexF64, dev.M), zeros(Float64, dev.M))
end

"""execute_m2_backward!(bplan, F_spec) (src/krfft.jl:1269-1277): returns
bplan.f0_buf (complex; its imaginary part is zero for the symmetric fields the
spectral ASU describes), overwritten by the next call."""
function execute_m2_backward!(bplan::M2BackwardPlanHIP, F_spec::AbstractVector{ComplexF64})
    _backward!(bplan.dev, F_spec, bplan.f0_real)
    bplan.f0_buf .= complex.(bplan.f0_real)
    return bplan.f0_buf
end

# ---------------------------------------------------------------------------
# Diffusion solver interface: src/diffusion_solver.jl
# ---------------------------------------------------------------------------

const _GPU_METHODS = Dict(:qfused => PATH_M2Q, :m2 => PATH_M2, :centered => PATH_M7, :qs => PATH_QS,
                          :auto => PATH_AUTO)

"A diffusion solver whose operator runs on the GPU (subgrid fields, M = N ./ L)."
struct GPUDiffusionSolver <: DiffusionSolver.AbstractDiffusionSolver
    dev::DevicePlan
    f0::Array{Float64,3}   # subgrid work buffer (the reference's QFusedDiffusionSolver.f0)
    w0::Array{Float64,3}
end

"""plan_diffusion(N, lattice, sg_num, dim, Δs; method=:matrix) (src/diffusion_solver.jl:165-252).
:qfused (the reference's Q-fused subgrid solver) and the extra subgrid methods
:m2, :centered, :qs and :auto run on the GPU; :matrix and :krfft are the
reference's own CPU solvers (the dense oracle and the block solver act on
real-ASU values, not on the subgrid the device path works on)."""
function plan_diffusion(N::Tuple, lattice::AbstractMatrix, sg_num::Int, dim::Int, Δs::Float64;
                        method=:matrix)
    if method in (:matrix, :krfft)
        return DiffusionSolver.plan_diffusion(N, lattice, sg_num, dim, Δs; method=method)
    end
    haskey(_GPU_METHODS, method) ||
        error("Unknown method: $method. Supported: :matrix, :krfft, :qfused, :m2, :centered, :qs, :auto")
    @assert length(N) == dim == 3
    dev = DevicePlan(NTuple{3,Int}(N), _ops_for(N, sg_num, dim), Δs, lattice; path=_GPU_METHODS[method])
    return GPUDiffusionSolver(dev, zeros(Float64, dev.M), zeros(Float64, dev.M))
end

"apply_diffusion_operator!(solver, u_subgrid) (src/diffusion_solver.jl:87-91)."
function DiffusionSolver.apply_diffusion_operator!(solver::GPUDiffusionSolver, u::AbstractArray{<:Real})
    _check_f0(solver.dev, u)
    copyto!(solver.f0, u)
    _diffuse!(solver.dev, solver.f0)
    copyto!(u, solver.f0)
    return u
end
apply_diffusion_operator!(solver, u) = DiffusionSolver.apply_diffusion_operator!(solver, u)

"""step_diffusion!(solver, u, w, dt) (src/diffusion_solver.jl:254-267):
u <- e^{-w dt/2} D e^{-w dt/2} u on the subgrid, in place.  (The reference's
interaction_step! has vector and ASU methods only; subgrid arrays of any shape
with M^3 entries are accepted here.)"""
function DiffusionSolver.step_diffusion!(solver::GPUDiffusionSolver, u::AbstractArray{<:Real},
                                         w::AbstractArray{<:Real}, dt::Float64)
    _check_f0(solver.dev, u)
    _check_f0(solver.dev, w)
    copyto!(solver.f0, u)
    copyto!(solver.w0, w)
    check(ccall((:cfft_step, libcfft), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Float64, Int32),
                solver.dev.handle, solver.f0, solver.w0, dt, Int32(1)))
    copyto!(u, solver.f0)
    return u
end
step_diffusion!(solver, u, w, dt::Float64) = DiffusionSolver.step_diffusion!(solver, u, w, dt)

# ---------------------------------------------------------------------------
# Block CFFT of a packed real-space ASU: src/cfft_plan.jl, src/execution.jl
# ---------------------------------------------------------------------------

mutable struct CFFTPlanHIP{D,T,A}
    asu::CrystallographicASU{D,T,A}
    handle::Ptr{Cvoid}
    order::Vector{Tuple{Int,Int}}   # (dimension group, block index) in device order
    total::Int
end

function _block_table(asu::CrystallographicASU{D}) where {D}
    order = Tuple{Int,Int}[]
    tab = Int32[]
    for d in sort!(collect(keys(asu.dim_blocks)))
        for (i, b) in enumerate(asu.dim_blocks[d])
            push!(order, (d, i))
            st = zeros(Int32, 9)
            for k in 1:D
                st[k] = Int32(first(b.range[k]))
                st[3 + k] = Int32(size(b.data, k))
                st[6 + k] = Int32(b.depth[k])
            end
            for k in D+1:3
                st[3 + k] = 1
            end
            append!(tab, st)
        end
    end
    return order, tab
end

"plan_cfft(N, sg_num, T=Float64, ArrayType=Array; dim=length(N)) (src/cfft_plan.jl:15-40)."
function plan_cfft(N::Tuple, sg_num::Int, T::Type=Float64, ArrayType=Array; dim::Int=length(N))
    points, shift = calc_asu(sg_num, dim, N)
    asu = pack_asu(points, N, T, ArrayType; shift=shift)
    order, tab = _block_table(asu)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:cfft_block_plan_create, libcfft), Cint, (Int32, Ptr{Int32}, Int64, Ptr{Ptr{Cvoid}}),
                Int32(dim), tab, Int64(length(order)), h))
    total = Ref{Int64}(0)
    nb = Ref{Int64}(0)
    check(ccall((:cfft_block_plan_info, libcfft), Cint, (Ptr{Cvoid}, Ptr{Int64}, Ptr{Int64}, Ptr{Int32}),
                h[], nb, total, C_NULL))
    p = CFFTPlanHIP(asu, h[], order, Int(total[]))
    finalizer(p) do x
        x.handle == C_NULL || ccall((:cfft_block_plan_destroy, libcfft), Cint, (Ptr{Cvoid},), x.handle)
        x.handle = C_NULL
    end
    return p
end

function _gather(p::CFFTPlanHIP, x::CrystallographicASU)
    buf = Vector{ComplexF64}(undef, p.total)
    o = 0
    for (d, i) in p.order
        haskey(x.dim_blocks, d) || error("Output ASU missing dimension $d blocks")
        v = x.dim_blocks[d][i].data
        n = length(v)
        copyto!(buf, o + 1, ComplexF64.(vec(v)), 1, n)
        o += n
    end
    return buf
end

function _scatter!(y::CrystallographicASU, p::CFFTPlanHIP, buf::Vector{ComplexF64})
    o = 0
    for (d, i) in p.order
        haskey(y.dim_blocks, d) || error("Output ASU missing dimension $d blocks")
        v = y.dim_blocks[d][i].data
        n = length(v)
        v .= reshape(view(buf, o+1:o+n), size(v))
        o += n
    end
    return y
end

"mul!(y, p, x) (src/execution.jl:10-31): unnormalised forward DFT of every block."
function mul!(y::CrystallographicASU, p::CFFTPlanHIP, x::CrystallographicASU)
    inb = _gather(p, x)
    out = similar(inb)
    check(ccall((:cfft_block_forward, libcfft), Cint, (Ptr{Cvoid}, Ptr{ComplexF64}, Ptr{ComplexF64}),
                p.handle, inb, out))
    return _scatter!(y, p, out)
end

"ldiv!(y, p, x) (src/execution.jl:37-64): normalised inverse DFT of every block."
function ldiv!(y::CrystallographicASU, p::CFFTPlanHIP, x::CrystallographicASU)
    inb = _gather(p, x)
    out = similar(inb)
    check(ccall((:cfft_block_inverse, libcfft), Cint, (Ptr{Cvoid}, Ptr{ComplexF64}, Ptr{ComplexF64}),
                p.handle, inb, out))
    return _scatter!(y, p, out)
end

end # module

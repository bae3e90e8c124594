# Julia binding of libcfft_hip.so for CrystallographicFFT.jl users.
#
# A maintainer adds this module beside the reference package; it keeps the
# reference's plan/execute names and argument meanings and forwards to the C
# ABI declared in include/cfft_hip.h.  Julia is not installed in the build
# image, so this file is written to the ABI but has not been run.
module CrystallographicFFTHIP

using LinearAlgebra
using CrystallographicFFT.SymmetryOps: SymOp, get_ops
using CrystallographicFFT.ASU: find_optimal_shift

const libcfft = get(ENV, "CFFT_HIP_LIB", joinpath(@__DIR__, "..", "libcfft_hip.so"))

const PATH_AUTO, PATH_M2, PATH_M7, PATH_M2Q, PATH_M7Q, PATH_QR, PATH_QS = Int32.(0:6)
const FP64, MIXED = Int32(0), Int32(1)

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
function pack_ops(ops::Vector{<:SymOp{3}})
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

mutable struct DevicePlan
    handle::Ptr{Cvoid}
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
    p = DevicePlan(h[], (info[1], info[2], info[3]), (info[4], info[5], info[6]),
                   Int(info[7]), Int(info[13]))
    finalizer(p) do x
        x.handle == C_NULL || ccall((:cfft_plan_destroy, libcfft), Cint, (Ptr{Cvoid},), x.handle)
        x.handle = C_NULL
    end
    return p
end

# --- reference-shaped API -------------------------------------------------

struct M2SCFTPlan
    dev::DevicePlan
end

"plan_m2_scft (src/q_fused_krfft.jl:948-986) on the GPU."
function plan_m2_scft(N::Tuple, sg_num::Int, dim::Int, Δs::Float64, lattice::AbstractMatrix)
    @assert length(N) == dim == 3
    _, ops_s = find_optimal_shift(get_ops(sg_num, dim, N), N)
    return M2SCFTPlan(DevicePlan(NTuple{3,Int}(N), ops_s, Δs, lattice; path=PATH_M2))
end

"execute_m2_scft! (src/q_fused_krfft.jl:989-1026): f0 <- D f0 in place."
function execute_m2_scft!(plan::M2SCFTPlan, f0::Array{Float64,3})
    size(f0) == plan.dev.M || error("f0 size $(size(f0)) does not match plan subgrid dims $(plan.dev.M).")
    check(ccall((:cfft_diffuse, libcfft), Cint, (Ptr{Cvoid}, Ptr{Float64}), plan.dev.handle, f0))
    return f0
end

"update_m2_kernel! (src/q_fused_krfft.jl:1036-1052)."
function update_m2_kernel!(plan::M2SCFTPlan, N, sg_num, dim, Δs_new::Float64, lattice)
    check(ccall((:cfft_update_kernel, libcfft), Cint, (Ptr{Cvoid}, Float64, Ptr{Float64}),
                plan.dev.handle, Δs_new, Float64.(vec(Matrix(lattice)))))
end

struct CenteredSCFTPlan
    dev::DevicePlan
end

"plan_centered_scft (src/centering_fold.jl:868-894)."
plan_centered_scft(spec, ops_shifted, N::Tuple, Δs::Float64, lattice::AbstractMatrix) =
    CenteredSCFTPlan(DevicePlan(NTuple{3,Int}(N), ops_shifted, Δs, lattice; path=PATH_M7))

"execute_centered_scft! (src/centering_fold.jl:908-931)."
function execute_centered_scft!(plan::CenteredSCFTPlan, f0::Array{Float64,3})
    check(ccall((:cfft_diffuse, libcfft), Cint, (Ptr{Cvoid}, Ptr{Float64}), plan.dev.handle, f0))
    return f0
end

"fft_reconstruct! (src/krfft.jl:554-562): real subgrid f0 -> spectral ASU."
function fft_reconstruct(dev::DevicePlan, f0::Array{Float64,3})
    F = Vector{ComplexF64}(undef, dev.n_spec * dev.batch)
    check(ccall((:cfft_forward, libcfft), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{ComplexF64}),
                dev.handle, f0, F))
    return F
end

"execute_m2_backward! / execute_centered_ikrfft!: spectral ASU -> f0."
function execute_backward!(dev::DevicePlan, F::Vector{ComplexF64}, f0::Array{Float64,3})
    check(ccall((:cfft_backward, libcfft), Cint, (Ptr{Cvoid}, Ptr{ComplexF64}, Ptr{Float64}),
                dev.handle, F, f0))
    return f0
end

"step_diffusion! (src/diffusion_solver.jl:254-267) with nsteps repetitions."
function step_diffusion!(dev::DevicePlan, q::Array{Float64,3}, w::Array{Float64,3}, dt::Float64;
                         nsteps::Int=1)
    check(ccall((:cfft_step, libcfft), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Float64, Int32),
                dev.handle, q, w, dt, Int32(nsteps)))
    return q
end

end # module

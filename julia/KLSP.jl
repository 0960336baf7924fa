# julia/KLSP.jl -- the reference's device-side API on libklsp (include/klsp.h).
#
# A drop-in for the GPU half of BasisFunctions.to_device, MatrixBuilder
# (compute_determinant, compute_determinant_batched, build_M_matrix),
# GridScan.grid_scan / find_minimum and NewtonSolver.newton_search
# (src/KalnajsLogSpiral/*.jl), with the same names, argument meanings and
# return types.  It is included by the package facade after the reference's
# submodules (src/KalnajsLogSpiral.jl), whose re-exports then bind to these
# functions:
#
#     include("KalnajsLogSpiral/KLSP.jl")
#     import .KLSP: DevicePrecomputed, to_device, compute_determinant,
#                   compute_determinant_batched, build_M_matrix, grid_scan,
#                   find_minimum, newton_search, NewtonResult
#
# so callers (run_kalnajs_gpu.jl:284, :307) are unchanged.  gpu_type == NONE,
# or devprecomp === nothing, routes to the reference's own CPU code, as the
# reference does (MatrixBuilder.jl:200-202); any other gpu_type runs on the
# gfx950 device through libklsp (there is no vendor dispatch: libklsp is
# HIP-only and fails with a status where no such device is present).
#
# Every ccall below takes the ABI's C types in the order klsp.h declares
# them; arrays cross as Julia's own column-major storage (no copies), the
# library keeps no host pointer after a call, and a non-zero status becomes
# an `error` with klsp_last_error's message.
module KLSP

using ..Configuration
using ..GPUBackend
using ..BasisFunctions
using ..MatrixBuilder
using ..GridScan
using ..NewtonSolver

export DevicePrecomputed, to_device, compute_determinant, compute_determinant_batched,
       build_M_matrix, grid_scan, find_minimum, newton_search, NewtonResult

const lib = get(ENV, "KLSP_LIB", "libklsp.so")

# klsp_precision
const PREC_FP64 = Cint(0)
const PREC_FP32 = Cint(1)
const PREC_MIXED = Cint(2)
const PREC_BF16X3 = Cint(3)
const PREC_MIXED_FP32 = Cint(4)
const PRECISIONS = Dict(:fp64 => PREC_FP64, :fp32 => PREC_FP32, :mixed => PREC_MIXED,
                        :bf16x3 => PREC_BF16X3, :mixed_fp32 => PREC_MIXED_FP32)

# ---------------------------------------------------------------------------
# C structs (field order and C types as in klsp.h)
# ---------------------------------------------------------------------------
struct Problem                      # klsp_problem
    NPh::Int64
    N_alpha::Int64
    n_l::Int64
    m::Cint
    G::Cdouble
    DJ::Ptr{Cdouble}
    F0l::Ptr{Cdouble}
    Omega_res::Ptr{Cdouble}
    N_kernel::Ptr{Cdouble}
end

struct NewtonParams                 # klsp_newton_params
    max_iter::Cint
    tol::Cdouble
    delta::Cdouble
    use_line_search::Cint
    max_line_search::Cint
    Omega_p_init::Cdouble
    gamma_init::Cdouble
    precision::Cint
end

mutable struct CNewtonResult        # klsp_newton_result
    Omega_p::Cdouble
    gamma::Cdouble
    omega::NTuple{2,Cdouble}
    det::NTuple{2,Cdouble}
    iterations::Cint
    converged::Cint
    n_det_evals::Cint
    n_fp64_builds::Cint
    n_lp_builds::Cint
    CNewtonResult() = new(0.0, 0.0, (0.0, 0.0), (0.0, 0.0), 0, 0, 0, 0, 0)
end

struct ScanParams                   # klsp_scan_params
    n_Op::Cint
    n_gamma::Cint
    Omega_p_min::Cdouble
    Omega_p_max::Cdouble
    gamma_min::Cdouble
    gamma_max::Cdouble
    precision::Cint
end

# ---------------------------------------------------------------------------
# Context: one HIP stream and the device-resident problem (klsp_create /
# klsp_destroy, the latter from the finalizer)
# ---------------------------------------------------------------------------
mutable struct Context
    ptr::Ptr{Cvoid}
    function Context(device::Integer=0)
        r = Ref{Ptr{Cvoid}}(C_NULL)
        status = ccall((:klsp_create, lib), Cint, (Cint, Ptr{Ptr{Cvoid}}), device, r)
        status == 0 || error("libklsp: cannot create a context on device $device (status $status)")
        c = new(r[])
        finalizer(c) do x
            x.ptr == C_NULL || ccall((:klsp_destroy, lib), Cint, (Ptr{Cvoid},), x.ptr)
            x.ptr = C_NULL
        end
        return c
    end
end

function check(ctx::Context, status::Cint)
    status == 0 && return nothing
    msg = unsafe_string(ccall((:klsp_last_error, lib), Cstring, (Ptr{Cvoid},), ctx.ptr))
    error("libklsp status $status: $msg")
end

"""
    DevicePrecomputed{T}

BasisFunctions.jl:381-395 with the same fields; `W_l_mat` is the libklsp
context that holds W (and its FP32 / BF16 copies) in HBM, and `precision` the
arithmetic of its M builds (FP64 for Tgpu = Float64, the complex64 / FP32-MFMA
build for Tgpu = Float32).  DJ_vec, F0l_all, Omega_res, N_kernel and d_alpha
are the host arrays, as in the reference (the context has its own device
copies).
"""
struct DevicePrecomputed{T<:AbstractFloat}
    m::Int
    G::T
    N_alpha::Int
    n_l::Int
    NPh::Int
    W_l_mat::Context
    DJ_vec::Vector{Float64}
    F0l_all::Matrix{Float64}
    Omega_res::Matrix{Float64}
    N_kernel::Vector{Float64}
    d_alpha::Vector{Float64}
    precision::Cint
end

"""
    to_device(precomp, gpu_type, Tgpu) -> DevicePrecomputed

BasisFunctions.jl:397-437: the omega-independent arrays and W_l_mat
[NPh, N_alpha, n_l] go to the device once (klsp_upload_problem,
klsp_upload_W), which then prepares the copy of W that `Tgpu` builds use
(klsp_prepare: the real basis of a symmetric alpha grid and the FP64 work
table, or the complex64 copy).  gpu_type == NONE returns the reference's host
copy.  `device` selects the HIP device (the reference selects it before the
call with gpu_device!).
"""
function to_device(precomp::PrecomputedData{T}, gpu_type, Tgpu::Type; device::Integer=0) where T
    gpu_type == GPUBackend.NONE && return BasisFunctions.to_device(precomp, gpu_type, Tgpu)
    ctx = Context(device)
    # convert copies only where the element type differs (W is 24.7 GB at the
    # default grid and crosses as it is)
    DJ = convert(Vector{Float64}, precomp.DJ_vec)
    F0l = convert(Matrix{Float64}, precomp.F0l_all)
    Ores = convert(Matrix{Float64}, precomp.Omega_res)
    Nk = convert(Vector{Float64}, precomp.N_kernel)
    W = convert(Array{ComplexF64,3}, precomp.W_l_mat)
    GC.@preserve DJ F0l Ores Nk W begin
        prob = Problem(precomp.NPh, precomp.N_alpha, precomp.n_l, precomp.m, precomp.G,
                       pointer(DJ), pointer(F0l), pointer(Ores), pointer(Nk))
        check(ctx, ccall((:klsp_upload_problem, lib), Cint, (Ptr{Cvoid}, Ref{Problem}),
                         ctx.ptr, prob))
        check(ctx, ccall((:klsp_upload_W, lib), Cint, (Ptr{Cvoid}, Ptr{ComplexF64}),
                         ctx.ptr, W))
    end
    prec = Tgpu == Float64 ? PREC_FP64 : PREC_FP32
    check(ctx, ccall((:klsp_prepare, lib), Cint, (Ptr{Cvoid}, Cint), ctx.ptr, prec))
    return DevicePrecomputed{Tgpu}(precomp.m, Tgpu(precomp.G), precomp.N_alpha, precomp.n_l,
                                   precomp.NPh, ctx, DJ, F0l, Ores, Nk,
                                   convert(Vector{Float64}, precomp.d_alpha), prec)
end

# ---------------------------------------------------------------------------
# MatrixBuilder
# ---------------------------------------------------------------------------
"""
    compute_determinant(omega, precomputed, devprecomp, gpu_type, Tcpu) -> Complex{Tcpu}

MatrixBuilder.jl:197-207: det(I - M(omega)) on the device (the weights, the
contraction, the G N epilogue and the FP64 complex LU with partial pivoting),
or the reference's CPU path for devprecomp === nothing / gpu_type == NONE.
"""
function compute_determinant(omega::Complex, precomputed::PrecomputedData,
                             devprecomp::Union{Nothing,DevicePrecomputed}, gpu_type,
                             Tcpu::Type)
    if devprecomp === nothing || gpu_type == GPUBackend.NONE
        return MatrixBuilder.compute_determinant(omega, precomputed, nothing,
                                                 GPUBackend.NONE, Tcpu)
    end
    return compute_determinant_batched([ComplexF64(omega)], precomputed, devprecomp,
                                       gpu_type, Tcpu)[1]
end

"""
    compute_determinant_batched(omegas, precomputed, devprecomp, gpu_type, Tcpu;
                                batch_size=100) -> Vector{Complex{Tcpu}}

MatrixBuilder.jl:215-239: one klsp_det call for every omega (the library
batches builds and LU launches itself; batch_size is accepted and unused).
"""
function compute_determinant_batched(omegas::AbstractVector{<:Complex},
                                     precomputed::PrecomputedData,
                                     devprecomp::Union{Nothing,DevicePrecomputed}, gpu_type,
                                     Tcpu::Type; batch_size::Int=100)
    if devprecomp === nothing || gpu_type == GPUBackend.NONE
        return MatrixBuilder.compute_determinant_batched(collect(omegas), precomputed, nothing,
                                                         GPUBackend.NONE, Tcpu;
                                                         batch_size=batch_size)
    end
    om = Vector{ComplexF64}(omegas)
    dets = Vector{ComplexF64}(undef, length(om))
    ctx = devprecomp.W_l_mat
    check(ctx, ccall((:klsp_det, lib), Cint,
                     (Ptr{Cvoid}, Ptr{ComplexF64}, Int64, Cint, Ptr{ComplexF64}),
                     ctx.ptr, om, length(om), devprecomp.precision, dets))
    return Vector{Complex{Tcpu}}(dets)
end

"""
    build_M_matrix(omega, precomputed) -> Matrix
    build_M_matrix(omega, devprecomp::DevicePrecomputed) -> Matrix{ComplexF64}

MatrixBuilder.jl:242-244 (the reference's debug entry, host build), and the
device build of the same M [N_alpha, N_alpha] (klsp_build_M).
"""
build_M_matrix(omega::Complex, precomputed::PrecomputedData) =
    MatrixBuilder.build_M_matrix(omega, precomputed)

function build_M_matrix(omega::Complex, devprecomp::DevicePrecomputed)
    N = devprecomp.N_alpha
    M = Matrix{ComplexF64}(undef, N, N)
    ctx = devprecomp.W_l_mat
    check(ctx, ccall((:klsp_build_M, lib), Cint,
                     (Ptr{Cvoid}, Cdouble, Cdouble, Cint, Ptr{ComplexF64}),
                     ctx.ptr, real(omega), imag(omega), devprecomp.precision, M))
    return M
end

# ---------------------------------------------------------------------------
# GridScan
# ---------------------------------------------------------------------------
"""
    grid_scan(config, precomputed, devprecomp, gpu_type, Tcpu; verbose=false,
              gpu_devices=Int[], precision=nothing) -> (det_grid, Omega_p_init, gamma_init)

GridScan.jl:25-82: det_grid [n_gamma, n_Op] over range(Omega_p_min,
Omega_p_max, n_Op) x range(gamma_min, gamma_max, n_gamma) and the argmin |det|
(find_minimum, GridScan.jl:234), in one klsp_grid_scan call.  One device:
gpu_devices beyond the first are ignored.  precision (:fp64, :fp32,
:bf16x3) overrides the device bundle's build arithmetic.
"""
function grid_scan(config::KalnajsConfig, precomputed::PrecomputedData,
                   devprecomp::Union{Nothing,DevicePrecomputed}, gpu_type, Tcpu::Type;
                   verbose::Bool=false, gpu_devices::Vector{Int}=Int[], precision=nothing)
    if devprecomp === nothing || gpu_type == GPUBackend.NONE
        return GridScan.grid_scan(config, precomputed, nothing, GPUBackend.NONE, Tcpu;
                                  verbose=verbose)
    end
    s = config.scan
    prec = precision === nothing ? devprecomp.precision : PRECISIONS[precision]
    sp = ScanParams(s.n_Op, s.n_gamma, s.Omega_p_min, s.Omega_p_max, s.gamma_min, s.gamma_max,
                    prec)
    det_grid = Matrix{ComplexF64}(undef, s.n_gamma, s.n_Op)
    Op0 = Ref{Cdouble}(0.0)
    g0 = Ref{Cdouble}(0.0)
    ctx = devprecomp.W_l_mat
    check(ctx, ccall((:klsp_grid_scan, lib), Cint,
                     (Ptr{Cvoid}, Ref{ScanParams}, Ptr{ComplexF64}, Ref{Cdouble}, Ref{Cdouble}),
                     ctx.ptr, sp, det_grid, Op0, g0))
    if verbose
        println("Minimum |det| = $(minimum(abs.(det_grid))) at Ω_p = $(Op0[]), γ = $(g0[])")
    end
    return Matrix{Complex{Tcpu}}(det_grid), Tcpu(Op0[]), Tcpu(g0[])
end

find_minimum(det_grid, Omega_p_grid, gamma_grid) =
    GridScan.find_minimum(det_grid, Omega_p_grid, gamma_grid)

# ---------------------------------------------------------------------------
# NewtonSolver
# ---------------------------------------------------------------------------
"""
    NewtonResult{T}

NewtonSolver.jl:19-27, field for field.
"""
struct NewtonResult{T<:AbstractFloat}
    Omega_p::T
    gamma::T
    omega::Complex{T}
    det_val::Complex{T}
    iterations::Int
    converged::Bool
    history::Matrix{T}   # [iteration, (Omega_p, gamma, |det|)]
end

"""
    newton_search(config, precomputed, devprecomp, gpu_type, Tcpu, Omega_p_init,
                  gamma_init; verbose=false, precision=nothing) -> NewtonResult{Tcpu}

NewtonSolver.jl:36-172 (forward-difference Newton on (Omega_p, gamma), the
optional halving line search, the bounds gamma >= 1e-4 and Omega_p in [0, 1],
the final re-evaluation) run by klsp_newton with every det on the device.
precision = :mixed selects the mixed-precision search (no counterpart in the
reference; klsp.h documents it: an FP64 root to |d omega| <= 1e-10 by a
calibrated estimate of its error, which from far starts may differ from the
root the FP64 iteration reaches).
"""
function newton_search(config::KalnajsConfig, precomputed::PrecomputedData,
                       devprecomp::Union{Nothing,DevicePrecomputed}, gpu_type, Tcpu::Type,
                       Omega_p_init::Real, gamma_init::Real; verbose::Bool=false,
                       precision=nothing)
    if devprecomp === nothing || gpu_type == GPUBackend.NONE
        r = NewtonSolver.newton_search(config, precomputed, nothing, GPUBackend.NONE, Tcpu,
                                       Omega_p_init, gamma_init; verbose=verbose)
        return NewtonResult{Tcpu}(r.Omega_p, r.gamma, r.omega, r.det_val, r.iterations,
                                  r.converged, r.history)
    end
    n = config.newton
    prec = precision === nothing ? devprecomp.precision : PRECISIONS[precision]
    np = NewtonParams(n.max_iter, n.tol, n.delta, n.use_line_search, n.max_line_search,
                      Omega_p_init, gamma_init, prec)
    res = CNewtonResult()
    hist = Matrix{Float64}(undef, max(n.max_iter, 1), 3)
    ctx = devprecomp.W_l_mat
    check(ctx, ccall((:klsp_newton, lib), Cint,
                     (Ptr{Cvoid}, Ref{NewtonParams}, Ref{CNewtonResult}, Ptr{Cdouble}),
                     ctx.ptr, np, res, hist))
    k = min(Int(res.iterations), max(n.max_iter, 1))
    if verbose
        for i in 1:k
            println("Iter $i: Ω_p=$(hist[i, 1])  γ=$(hist[i, 2])  |det|=$(hist[i, 3])")
        end
        println(res.converged != 0 ? "✓ Converged!" : "Not converged")
    end
    return NewtonResult{Tcpu}(Tcpu(res.Omega_p), Tcpu(res.gamma),
                              Complex{Tcpu}(res.omega[1], res.omega[2]),
                              Complex{Tcpu}(res.det[1], res.det[2]), Int(res.iterations),
                              res.converged != 0, Matrix{Tcpu}(hist[1:k, :]))
end

end # module KLSP

# HIPExt: FluxOptics.jl device backend over libfluxhip (MI355X, gfx950).
#
# This file is the Julia side of the drop-in: a package extension that FluxOptics.jl would load
# next to (or instead of) its CUDA extension (src/FluxOptics.jl:104-115 loads CUDAExt through
# Requires).  It dispatches on ScalarField{<:ROCArray} and calls libfluxhip through ccall on raw
# device pointers.  Julia is not installed in the build image, so this file has not been run;
# the same C ABI is exercised by the Python host layer and the GPU tests.

module HIPExt

using AMDGPU: ROCArray, AMDGPU
using ..FluxOptics: ScalarField, ASProp, RSProp, Phase, AS_BPM, Forward, Backward, Direction
import ..FluxOptics: propagate!, backpropagate!, propagate_and_save!, backpropagate_with_gradient!
import ChainRulesCore

const libfluxhip = get(ENV, "FLUXHIP_LIB", joinpath(@__DIR__, "..", "fluxoptics", "libfluxhip.so"))

const CtxPtr = Ptr{Cvoid}
const PlanPtr = Ptr{Cvoid}

function check(rc::Cint, fn::Symbol)
    rc == 0 && return nothing
    msg = unsafe_string(ccall((:fh_last_error, libfluxhip), Cstring, ()))
    error("libfluxhip $(fn) failed ($(rc)): $(msg)")
end

# one context per device, bound to the task-local HIP stream
const CONTEXTS = Dict{Int, CtxPtr}()
function context()
    dev = AMDGPU.device_id(AMDGPU.device()) - 1
    ctx = get!(CONTEXTS, dev) do
        r = Ref{CtxPtr}(C_NULL)
        check(ccall((:fh_ctx_create, libfluxhip), Cint, (Cint, Ptr{Cvoid}, Ptr{CtxPtr}), dev, C_NULL, r), :fh_ctx_create)
        r[]
    end
    stream = AMDGPU.stream().stream
    check(ccall((:fh_ctx_set_stream, libfluxhip), Cint, (CtxPtr, Ptr{Cvoid}), ctx, stream), :fh_ctx_set_stream)
    ctx
end

dir_sign(::Type{Forward}) = Cint(1)
dir_sign(::Type{Backward}) = Cint(-1)

# lambdas as the reference stores them (Float32 for ComplexF32 fields, Fields.jl:28-32), promoted
lambdas_f64(u::ScalarField) = Float64.(collect(vec(u.lambdas.collection)))

# plan cache keyed by (propagator object, batch, lambdas)
const PLANS = IdDict{Any, Dict{Any, PlanPtr}}()

function asprop_plan(p, u::ScalarField{<:ROCArray{ComplexF32}}, z::Float64, n0::Float64)
    nx, ny = size(u.electric, 1), size(u.electric, 2)
    batch = div(length(u.electric), nx * ny)
    lams = lambdas_f64(u)
    cache = get!(PLANS, p) do; Dict{Any, PlanPtr}() end
    get!(cache, (batch, lams)) do
        r = Ref{PlanPtr}(C_NULL)
        check(ccall((:fh_asprop_create, libfluxhip), Cint,
                    (CtxPtr, Cint, Cint, Cint, Cdouble, Cdouble, Ptr{Cdouble}, Cint, Cdouble, Cdouble, Ptr{PlanPtr}),
                    context(), nx, ny, batch, u.ds[1], u.ds[2], lams, length(lams), z, n0, r), :fh_asprop_create)
        r[]
    end
end

function rsprop_plan(p, u::ScalarField{<:ROCArray{ComplexF32}}, z::Float64)
    nx, ny = size(u.electric, 1), size(u.electric, 2)
    batch = div(length(u.electric), nx * ny)
    lams = lambdas_f64(u)
    cache = get!(PLANS, p) do; Dict{Any, PlanPtr}() end
    get!(cache, (batch, lams)) do
        r = Ref{PlanPtr}(C_NULL)
        check(ccall((:fh_rsprop_create, libfluxhip), Cint,
                    (CtxPtr, Cint, Cint, Cint, Cdouble, Cdouble, Ptr{Cdouble}, Cint, Cdouble, Ptr{PlanPtr}),
                    context(), nx, ny, batch, u.ds[1], u.ds[2], lams, length(lams), z, r), :fh_rsprop_create)
        r[]
    end
end

# ASProp / RSProp are AbstractSequences (ft, kernel, ift[, pad/crop]); the whole sequence runs
# as one libfluxhip propagation.  z and n0 come from the kernel component (angular_spectrum.jl:78-102).
function propagate!(u::ScalarField{<:ROCArray{ComplexF32}}, p::ASProp, direction::Type{<:Direction})
    k = p.optical_components[2]
    plan = asprop_plan(p, u, Float64(k.z), Float64(k.n0))
    check(ccall((:fh_propagate, libfluxhip), Cint, (PlanPtr, Ptr{Cvoid}, Cint), plan, pointer(u.electric), dir_sign(direction)), :fh_propagate)
    u
end

function propagate!(u::ScalarField{<:ROCArray{ComplexF32}}, p::RSProp, direction::Type{<:Direction})
    k = p.optical_components[3]
    plan = rsprop_plan(p, u, Float64(k.z))
    check(ccall((:fh_propagate, libfluxhip), Cint, (PlanPtr, Ptr{Cvoid}, Cint), plan, pointer(u.electric), dir_sign(direction)), :fh_propagate)
    u
end

backpropagate!(u::ScalarField{<:ROCArray{ComplexF32}}, p::Union{ASProp, RSProp}, direction::Type{<:Direction}) =
    propagate!(u, p, reverse(direction))

# Phase (phasemask.jl:90-150)
function propagate!(u::ScalarField{<:ROCArray{ComplexF32}}, p::Phase, direction::Type{<:Direction})
    nx, ny = size(u.electric, 1), size(u.electric, 2)
    batch = div(length(u.electric), nx * ny)
    check(ccall((:fh_phase_apply, libfluxhip), Cint,
                (CtxPtr, Ptr{Cvoid}, Ptr{Cfloat}, Cint, Cint, Cint, Cint, Ptr{Cvoid}),
                context(), pointer(u.electric), pointer(p.ϕ), nx, ny, batch, dir_sign(direction), C_NULL), :fh_phase_apply)
    u
end

function propagate_and_save!(u::ScalarField{<:ROCArray{ComplexF32}}, u_saved::ROCArray, p::Phase, direction::Type{<:Direction})
    nx, ny = size(u.electric, 1), size(u.electric, 2)
    batch = div(length(u.electric), nx * ny)
    check(ccall((:fh_phase_apply, libfluxhip), Cint,
                (CtxPtr, Ptr{Cvoid}, Ptr{Cfloat}, Cint, Cint, Cint, Cint, Ptr{Cvoid}),
                context(), pointer(u.electric), pointer(p.ϕ), nx, ny, batch, dir_sign(direction), pointer(u_saved)), :fh_phase_apply)
    u
end

function backpropagate_with_gradient!(∂v::ScalarField{<:ROCArray{ComplexF32}}, u_saved::ROCArray,
                                      ∂p::NamedTuple, p::Phase, direction::Type{<:Direction})
    nx, ny = size(∂v.electric, 1), size(∂v.electric, 2)
    batch = div(length(∂v.electric), nx * ny)
    check(ccall((:fh_phase_backward, libfluxhip), Cint,
                (CtxPtr, Ptr{Cvoid}, Ptr{Cfloat}, Ptr{Cvoid}, Ptr{Cfloat}, Cint, Cint, Cint, Cint, Cint),
                context(), pointer(∂v.electric), pointer(p.ϕ), pointer(u_saved), pointer(∂p.ϕ),
                nx, ny, batch, dir_sign(direction), 0), :fh_phase_backward)
    (∂v, ∂p)
end

# The rrules of ChainRulesCoreExt.jl:21-111 call propagate!/backpropagate! and
# backpropagate_with_gradient!, so the methods above are picked up by dispatch; the fused
# Phase→propagator pair (fh_phase_propagate / fh_propagate_phase_backward) is used by a
# merge rule on OpticalSequence (merge_rules.jl) when a trainable Phase precedes an ASProp.

end # module

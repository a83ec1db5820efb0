# HIPExt: FluxOptics.jl device backend over libfluxhip (MI355X, gfx950).
#
# This file is the Julia side of the drop-in: a package extension that FluxOptics.jl would load
# next to (or instead of) its CUDA extension (src/FluxOptics.jl:104-115 loads CUDAExt through
# Requires).  It dispatches on ScalarField{<:ROCArray} and calls libfluxhip through ccall on raw
# device pointers.  Julia is not installed in the build image, so this file has not been run;
# the same C ABI is exercised by the Python host layer and the GPU tests.

module HIPExt

using AMDGPU: ROCArray, AMDGPU
using ..FluxOptics: ScalarField, ASProp, RSProp, ShiftProp, CollinsProp, Phase, AS_BPM, Forward, Backward, Direction,
                    set_field_ds!
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

# fh_prop_desc of include/fluxhip.h (C layout: Julia's struct layout rules match the C ABI)
struct PropDesc
    nx::Cint; ny::Cint; batch::Cint
    dx::Cdouble; dy::Cdouble
    lambdas::Ptr{Cdouble}; nlambdas::Cint
    z::Ptr{Cdouble}; nz::Cint
    tilts_x::Ptr{Cdouble}; tilts_y::Ptr{Cdouble}; ntilts::Cint
    n0::Cdouble
    flags::Cint
end
const FH_PROP_PARAXIAL = Cint(1)
const FH_PROP_TRACK_TILTS = Cint(2)
const FH_PROP_SHIFT = Cint(4)

# tilts as the reference stores them (one collection per axis, Fields.jl:34-37), promoted
tilts_f64(u::ScalarField, batch) = map(c -> (v = Float64.(collect(vec(c))); length(v) == 1 ? fill(v[1], batch) : v),
                                       u.tilts.collection)

# plan cache keyed by (propagator object, batch, lambdas, tilts)
const PLANS = IdDict{Any, Dict{Any, PlanPtr}}()

function prop_plan(p, u::ScalarField{<:ROCArray{ComplexF32}}, z::Float64, n0::Float64, flags::Cint, fn::Symbol)
    nx, ny = size(u.electric, 1), size(u.electric, 2)
    batch = div(length(u.electric), nx * ny)
    lams = lambdas_f64(u)
    tx, ty = tilts_f64(u, batch)
    cache = get!(PLANS, p) do; Dict{Any, PlanPtr}() end
    get!(cache, (batch, lams, tx, ty)) do
        zs = [z]
        r = Ref{PlanPtr}(C_NULL)
        GC.@preserve lams zs tx ty begin
            d = Ref(PropDesc(nx, ny, batch, u.ds[1], u.ds[2], pointer(lams), length(lams), pointer(zs), 1,
                             pointer(tx), pointer(ty), batch, n0, flags))
            check(ccall((fn, libfluxhip), Cint, (CtxPtr, Ptr{PropDesc}, Ptr{PlanPtr}), context(), d, r), fn)
        end
        r[]
    end
end

# ASProp / RSProp are AbstractSequences (ft, kernel, ift[, pad/crop]); the whole sequence runs
# as one libfluxhip propagation.  z, n0 and the kernel options come from the kernel component
# (angular_spectrum.jl:78-102, rayleigh_sommerfeld.jl:38-60).
function propagate!(u::ScalarField{<:ROCArray{ComplexF32}}, p::ASProp, direction::Type{<:Direction})
    k = p.optical_components[2]
    k.filter === nothing || error("libfluxhip: ASProp filters are not on the accelerated path")
    flags = (k.is_paraxial ? FH_PROP_PARAXIAL : Cint(0)) | (k.track_tilts ? FH_PROP_TRACK_TILTS : Cint(0))
    plan = prop_plan(p, u, Float64(k.z), Float64(k.n0), flags, :fh_asprop_create_ex)
    check(ccall((:fh_propagate, libfluxhip), Cint, (PlanPtr, Ptr{Cvoid}, Cint), plan, pointer(u.electric), dir_sign(direction)), :fh_propagate)
    u
end

function propagate!(u::ScalarField{<:ROCArray{ComplexF32}}, p::RSProp, direction::Type{<:Direction})
    k = p.optical_components[3]
    flags = k.track_tilts ? FH_PROP_TRACK_TILTS : Cint(0)
    plan = prop_plan(p, u, Float64(k.z), 1.0, flags, :fh_rsprop_create_ex)
    check(ccall((:fh_propagate, libfluxhip), Cint, (PlanPtr, Ptr{Cvoid}, Cint), plan, pointer(u.electric), dir_sign(direction)), :fh_propagate)
    u
end

# ShiftProp (shift_kernel.jl:91-117): (ft, ShiftKernel, ift), z from the kernel component
function propagate!(u::ScalarField{<:ROCArray{ComplexF32}}, p::ShiftProp, direction::Type{<:Direction})
    k = p.optical_components[2]
    plan = prop_plan(p, u, Float64(k.z), 1.0, FH_PROP_SHIFT, :fh_asprop_create_ex)
    check(ccall((:fh_propagate, libfluxhip), Cint, (PlanPtr, Ptr{Cvoid}, Cint), plan, pointer(u.electric), dir_sign(direction)), :fh_propagate)
    u
end

# CollinsProp (collins_integral.jl:262-296): (a_chirp, pad, ft, conv, ift, crop, d_chirp) as one
# plan; A, B, ds from the A chirp, D and ds' from the D chirp; the field's sampling is relabelled
# as the chirp components do (set_field_ds!).
const COLLINS = IdDict{Any, Dict{Any, PlanPtr}}()
function propagate!(u::ScalarField{<:ROCArray{ComplexF32}}, p::CollinsProp, direction::Type{<:Direction})
    ac, dc = p.optical_components[1], p.optical_components[end]
    nx, ny = size(u.electric, 1), size(u.electric, 2)
    batch = div(length(u.electric), nx * ny)
    lams = lambdas_f64(u)
    tx, ty = tilts_f64(u, batch)
    ds, dsp = Float64.(ac.ds), Float64.(dc.ds′)
    cache = get!(COLLINS, p) do; Dict{Any, PlanPtr}() end
    plan = get!(cache, (batch, lams, tx, ty)) do
        r = Ref{PlanPtr}(C_NULL)
        GC.@preserve lams tx ty begin
            d = Ref(PropDesc(nx, ny, batch, ds[1], ds[2], pointer(lams), length(lams), C_NULL, 0,
                             pointer(tx), pointer(ty), batch, 1.0, ac.track_tilts ? FH_PROP_TRACK_TILTS : Cint(0)))
            check(ccall((:fh_collins_create, libfluxhip), Cint,
                        (CtxPtr, Ptr{PropDesc}, Cdouble, Cdouble, Cdouble, Cdouble, Cdouble, Ptr{PlanPtr}),
                        context(), d, dsp[1], dsp[2], Float64(ac.a), Float64(ac.b), Float64(dc.d), r), :fh_collins_create)
        end
        r[]
    end
    check(ccall((:fh_propagate, libfluxhip), Cint, (PlanPtr, Ptr{Cvoid}, Cint), plan, pointer(u.electric), dir_sign(direction)), :fh_propagate)
    set_field_ds!(u, direction == Forward ? dsp : ds)
    u
end

backpropagate!(u::ScalarField{<:ROCArray{ComplexF32}}, p::Union{ASProp, RSProp, ShiftProp, CollinsProp},
               direction::Type{<:Direction}) = propagate!(u, p, reverse(direction))

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

# SquaredIntensityDifference (squared_intensity_difference.jl:78-90) on the device: the values
# (FP64, one per field or one for a spatial-only target) and, in the pullback, 4 dc d u.
function sq_intensity_diff!(loss::ROCArray{Float64}, du::Union{Nothing, ROCArray{ComplexF32}},
                            u::ScalarField{<:ROCArray{ComplexF32}}, target::ROCArray{Float32},
                            dc::Union{Nothing, ROCArray{Float64}} = nothing)
    nx, ny = size(u.electric, 1), size(u.electric, 2)
    batch = div(length(u.electric), nx * ny)
    tb = div(length(target), nx * ny)
    check(ccall((:fh_sq_intensity_diff, libfluxhip), Cint,
                (CtxPtr, Ptr{Cvoid}, Ptr{Cfloat}, Cint, Cint, Cint, Cint, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cvoid}),
                context(), pointer(u.electric), pointer(target), nx, ny, batch, tb,
                dc === nothing ? C_NULL : pointer(dc), pointer(loss), du === nothing ? C_NULL : pointer(du)),
          :fh_sq_intensity_diff)
    loss
end

# The rrules of ChainRulesCoreExt.jl:21-111 call propagate!/backpropagate! and
# backpropagate_with_gradient!, so the methods above are picked up by dispatch; the fused
# Phase→propagator pair (fh_phase_propagate / fh_propagate_phase_backward) is used by a
# merge rule on OpticalSequence (merge_rules.jl) when a trainable Phase precedes an ASProp.

end # module

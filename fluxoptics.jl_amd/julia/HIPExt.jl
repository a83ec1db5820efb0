# HIPExt: FluxOptics.jl device backend over libfluxhip (MI355X, gfx950).
#
# This file is the Julia side of the drop-in: a package extension that FluxOptics.jl would load
# next to (or instead of) its CUDA extension (src/FluxOptics.jl:104-115 loads CUDAExt through
# Requires).  It dispatches on ScalarField{<:ROCArray{ComplexF32}, Nd} and calls libfluxhip through
# ccall on raw device pointers.  Julia is not installed in the build image, so this file has not
# been run; every ccall in it is checked against include/fluxhip.h by a CPU test (name, arity and
# argument types), and the same C ABI is exercised by the Python host layer and the GPU tests.

module HIPExt

using AMDGPU: ROCArray, AMDGPU
using LinearAlgebra: Adjoint
using ..FluxOptics: ScalarField, ASProp, RSProp, ShiftProp, CollinsProp, Phase, AS_BPM, Shift_BPM, Forward,
                    Backward, Direction, Trainable, Buffered, Unbuffered, set_field_ds!
using ..FluxOptics.OpticalComponents: BPM
import ..FluxOptics: propagate!, propagate, backpropagate!, backpropagate
import ..FluxOptics.OpticalComponents: propagate_and_save!, backpropagate_with_gradient!, alloc_saved_buffer
import ChainRulesCore
using ChainRulesCore: NoTangent, unthunk

const libfluxhip = get(ENV, "FLUXHIP_LIB", joinpath(@__DIR__, "..", "fluxoptics", "libfluxhip.so"))

const CtxPtr = Ptr{Cvoid}
const PlanPtr = Ptr{Cvoid}
const BpmPtr = Ptr{Cvoid}

# device fields of either dimensionality, and of each
const DField = ScalarField{<:ROCArray{ComplexF32}}
const DField2 = ScalarField{<:ROCArray{ComplexF32}, 2}
const DField1 = ScalarField{<:ROCArray{ComplexF32}, 1}

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

# The library's view of a field: (nx, ny, batch), ny = 1 for a 1-D field [nx, batch...]
# (Fields.jl:135-136, include/fluxhip.h: 1-D fields are passed with ny = 1).
function grid(u::ScalarField{<:Any, Nd}) where {Nd}
    nx = size(u.electric, 1)
    ny = Nd == 2 ? size(u.electric, 2) : 1
    (nx, ny, div(length(u.electric), nx * ny))
end

# lambdas as the reference stores them (Float32 for ComplexF32 fields, Fields.jl:28-32), promoted
lambdas_f64(u::ScalarField) = Float64.(collect(vec(u.lambdas.collection)))

# tilts as the reference stores them (one collection per axis, Fields.jl:34-37), promoted; a
# 1-D field has theta_y = 0
function tilts_f64(u::ScalarField{<:Any, Nd}, batch) where {Nd}
    t = map(c -> (v = Float64.(collect(vec(c))); length(v) == 1 ? fill(v[1], batch) : v), u.tilts.collection)
    Nd == 2 ? t : (t[1], zeros(batch))
end

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

# The sampling a propagator was built for (its kernel's construction-time ds, not the field's):
# FourierKernel stores fftfreq(n, 1/dx) (fourier_kernel.jl:17), whose second entry is 1/(n dx);
# ConvolutionKernel stores the offsets circshift((1-n):(n-1), n) .* dx (convolution_kernel.jl:16),
# whose second entry is dx.  Both hold the field's real type; the field's own ds is taken when
# the stored value is its rounding, so the common case keeps the exact Float64 sampling.
host_vec(a) = Float64.(collect(vec(a isa Adjoint ? parent(a) : a)))
function snap(d::Float64, ref::Real)
    abs(d - ref) <= 4 * eps(Float32) * abs(ref) ? Float64(ref) : d
end
function kernel_ds(k, u::ScalarField{<:Any, Nd}) where {Nd}
    if hasproperty(k.kernel, :f_vec)
        fv = k.kernel.f_vec
        dx = 1 / (length(host_vec(fv.x)) * host_vec(fv.x)[2])
        dy = Nd == 2 ? 1 / (length(host_vec(fv.y)) * host_vec(fv.y)[2]) : 1.0
    else
        sv = k.kernel.s_vec
        dx = host_vec(sv.x)[2]
        dy = Nd == 2 ? host_vec(sv.y)[2] : 1.0
    end
    (snap(dx, u.ds[1]), Nd == 2 ? snap(dy, u.ds[2]) : 1.0)
end

# The filter of an ASProp kernel sampled as its kernel calls it (angular_spectrum.jl:5,34,57 and
# the 1-D forms :18,40,71): on the fftfreq grid of ds, at (fx, fy) for the on-axis and paraxial
# kernels, at (fx + f0x, fy + f0y) per field for the tilted one; a 1-D filter takes fx only.
# Returns the interleaved (re, im) grids and their count (include/fluxhip.h,
# fh_asprop_create_filtered).
fftfreq64(n, d) = [(k < (n + 1) ÷ 2 ? k : k - n) / (n * d) for k in 0:(n - 1)]
function filter_grids(k, nx, ny, ds, lams, tx, ty, n0, Nd)
    fx, fy = fftfreq64(nx, ds[1]), Nd == 2 ? fftfreq64(ny, ds[2]) : [0.0]
    f(a, b) = ComplexF64(Nd == 2 ? k.filter(a, b) : k.filter(a))
    tilted = !k.is_paraxial && (any(!iszero, tx) || any(!iszero, ty))
    grids = if tilted
        [[f(fx[i] + sin(tx[b]) * n0 / lams[b], fy[j] + sin(ty[b]) * n0 / lams[b]) for i in 1:nx, j in 1:ny]
         for b in eachindex(tx)]
    else
        [[f(fx[i], fy[j]) for i in 1:nx, j in 1:ny]]
    end
    buf = Vector{Float64}(undef, 2 * nx * ny * length(grids))
    for (g, a) in enumerate(grids), (q, v) in enumerate(vec(a))   # x fastest: [j][i] of the header
        buf[2 * ((g - 1) * nx * ny + q) - 1] = real(v)
        buf[2 * ((g - 1) * nx * ny + q)] = imag(v)
    end
    buf, Cint(length(grids))
end

# plan cache keyed by (propagator object, device, batch, lambdas, tilts, ds)
const PLANS = IdDict{Any, Dict{Any, PlanPtr}}()

function prop_plan(p, k, u::ScalarField{<:Any, Nd}, z::Float64, n0::Float64, flags::Cint,
                   fn::Symbol) where {Nd}
    nx, ny, batch = grid(u)
    lams = lambdas_f64(u)
    tx, ty = tilts_f64(u, batch)
    ds = kernel_ds(k, u)
    cache = get!(PLANS, p) do; Dict{Any, PlanPtr}() end
    get!(cache, (AMDGPU.device_id(AMDGPU.device()), batch, lams, tx, ty, ds)) do
        zs = [z]
        r = Ref{PlanPtr}(C_NULL)
        GC.@preserve lams zs tx ty begin
            d = Ref(PropDesc(nx, ny, batch, ds[1], ds[2], pointer(lams), length(lams), pointer(zs), 1,
                             pointer(tx), pointer(ty), batch, n0, flags))
            filt = hasproperty(k, :filter) ? k.filter : nothing
            if filt === nothing
                check(ccall((fn, libfluxhip), Cint, (CtxPtr, Ptr{PropDesc}, Ptr{PlanPtr}), context(), d, r), fn)
            else
                buf, nf = filter_grids(k, nx, ny, ds, lams, tx, ty, n0, Nd)
                GC.@preserve buf begin
                    check(ccall((:fh_asprop_create_filtered, libfluxhip), Cint,
                                (CtxPtr, Ptr{PropDesc}, Ptr{Cdouble}, Cint, Ptr{PlanPtr}),
                                context(), d, pointer(buf), nf, r), :fh_asprop_create_filtered)
                end
            end
        end
        r[]
    end
end

run_plan(plan, u::DField, direction) =
    check(ccall((:fh_propagate, libfluxhip), Cint, (PlanPtr, Ptr{Cvoid}, Cint), plan, pointer(u.electric),
                dir_sign(direction)), :fh_propagate)

# ASProp / RSProp / ShiftProp are AbstractSequences (ft, kernel, ift[, pad/crop]); the whole
# sequence runs as one libfluxhip propagation.  z, n0, the kernel options and the sampling come
# from the kernel component (angular_spectrum.jl:78-102, rayleigh_sommerfeld.jl:38-60,
# shift_kernel.jl:14-30).
function propagate!(u::DField, p::ASProp, direction::Type{<:Direction})
    k = p.optical_components[2]
    flags = (k.is_paraxial ? FH_PROP_PARAXIAL : Cint(0)) | (k.track_tilts ? FH_PROP_TRACK_TILTS : Cint(0))
    run_plan(prop_plan(p, k, u, Float64(k.z), Float64(k.n0), flags, :fh_asprop_create_ex), u, direction)
    u
end

function propagate!(u::DField2, p::RSProp, direction::Type{<:Direction})
    k = p.optical_components[3]
    flags = k.track_tilts ? FH_PROP_TRACK_TILTS : Cint(0)
    run_plan(prop_plan(p, k, u, Float64(k.z), 1.0, flags, :fh_rsprop_create_ex), u, direction)
    u
end
propagate!(::DField1, ::RSProp, ::Type{<:Direction}) =
    error("libfluxhip: RSProp has no 1-D kernel (rs_kernel takes x and y)")

function propagate!(u::DField, p::ShiftProp, direction::Type{<:Direction})
    k = p.optical_components[2]
    run_plan(prop_plan(p, k, u, Float64(k.z), 1.0, FH_PROP_SHIFT, :fh_asprop_create_ex), u, direction)
    u
end

# CollinsProp (collins_integral.jl:262-296): (a_chirp, pad, ft, conv, ift, crop, d_chirp) as one
# plan; A, B, ds from the A chirp, D and ds' from the D chirp; the field's sampling is relabelled
# as the chirp components do (set_field_ds!).  2-D fields only.
const COLLINS = IdDict{Any, Dict{Any, PlanPtr}}()
function propagate!(u::DField2, p::CollinsProp, direction::Type{<:Direction})
    ac, dc = p.optical_components[1], p.optical_components[end]
    nx, ny, batch = grid(u)
    lams = lambdas_f64(u)
    tx, ty = tilts_f64(u, batch)
    ds, dsp = Float64.(ac.ds), Float64.(dc.ds′)
    cache = get!(COLLINS, p) do; Dict{Any, PlanPtr}() end
    plan = get!(cache, (AMDGPU.device_id(AMDGPU.device()), batch, lams, tx, ty)) do
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
    run_plan(plan, u, direction)
    set_field_ds!(u, direction == Forward ? dsp : ds)
    u
end
propagate!(::DField1, ::CollinsProp, ::Type{<:Direction}) = error("libfluxhip: CollinsProp takes 2-D fields")

const DeviceProp = Union{ASProp, RSProp, ShiftProp, CollinsProp}

# The adjoint of these propagators is the propagation in the reverse direction
# (freespace.jl:61-63; CollinsProp's Backward, collins_integral.jl:189, as in the reference).
backpropagate!(u::DField, p::DeviceProp, direction::Type{<:Direction}) = propagate!(u, p, reverse(direction))
propagate(u::DField, p::DeviceProp, direction::Type{<:Direction}) = propagate!(copy(u), p, direction)
backpropagate(u::DField, p::DeviceProp, direction::Type{<:Direction}) = backpropagate!(copy(u), p, direction)

# rrules of the accelerated sequences.  The reference differentiates ASProp / RSProp / ShiftProp /
# CollinsProp by Zygote tracing through their components (they are AbstractSequences, with rrules
# only for AbstractCustomComponent, ChainRulesCoreExt.jl:21-111); the methods above replace that
# body with a ccall, so the pullback is supplied here, in the form of ChainRulesCoreExt.jl:21-32
# and :67-78: the adjoint propagation of the cotangent.
function ChainRulesCore.rrule(::typeof(propagate!), u::DField, p::DeviceProp, direction::Type{<:Direction})
    v = propagate!(u, p, direction)
    pullback(∂v) = (NoTangent(), backpropagate!(unthunk(∂v), p, direction), NoTangent(), NoTangent())
    v, pullback
end

function ChainRulesCore.rrule(::typeof(propagate), u::DField, p::DeviceProp, direction::Type{<:Direction})
    v = propagate(u, p, direction)
    pullback(∂v) = (NoTangent(), backpropagate(unthunk(∂v), p, direction), NoTangent(), NoTangent())
    v, pullback
end

# Phase (phasemask.jl:90-150); a 1-D field's mask is a row of nx values.  Phase is an
# AbstractCustomComponent: the reference's own rrules (ChainRulesCoreExt.jl:21-111) dispatch to
# these methods.
function phase_apply!(u::DField, p::Phase, direction, saved)
    nx, ny, batch = grid(u)
    check(ccall((:fh_phase_apply, libfluxhip), Cint,
                (CtxPtr, Ptr{Cvoid}, Ptr{Cfloat}, Cint, Cint, Cint, Cint, Ptr{Cvoid}),
                context(), pointer(u.electric), pointer(p.ϕ), nx, ny, batch, dir_sign(direction), saved), :fh_phase_apply)
    u
end
propagate!(u::DField, p::Phase, direction::Type{<:Direction}) = phase_apply!(u, p, direction, C_NULL)
backpropagate!(u::DField, p::Phase, direction::Type{<:Direction}) = phase_apply!(u, p, reverse(direction), C_NULL)
propagate_and_save!(u::DField, u_saved::ROCArray, p::Phase, direction::Type{<:Direction}) =
    phase_apply!(u, p, direction, pointer(u_saved))

function backpropagate_with_gradient!(∂v::DField, u_saved::ROCArray, ∂p::NamedTuple, p::Phase,
                                      direction::Type{<:Direction})
    nx, ny, batch = grid(∂v)
    check(ccall((:fh_phase_backward, libfluxhip), Cint,
                (CtxPtr, Ptr{Cvoid}, Ptr{Cfloat}, Ptr{Cvoid}, Ptr{Cfloat}, Cint, Cint, Cint, Cint, Cint),
                context(), pointer(∂v.electric), pointer(p.ϕ), pointer(u_saved), pointer(∂p.ϕ),
                nx, ny, batch, dir_sign(direction), 0), :fh_phase_backward)
    (∂v, ∂p)
end

# AS_BPM / Shift_BPM (bpm.jl:8-74,244-312).  The library keeps one checkpoint every
# BPM_CHECKPOINT[] slices (2-D) or every slice input (1-D) inside its BPM object, so the
# reference's saved stack is not used: alloc_saved_buffer for an unbuffered trainable BPM returns
# an empty array (the reference would allocate the whole [nx, ny, batch, nslices] stack).  BPM is
# an AbstractCustomComponent: the reference's rrules dispatch to the methods below.
const BPM_CHECKPOINT = Ref(20)
const BPMS = IdDict{Any, Dict{Any, BpmPtr}}()

function bpm_obj(p::BPM, u::ScalarField{<:Any, Nd}) where {Nd}
    nx, ny, batch = grid(u)
    lams = lambdas_f64(u)
    tx, ty = tilts_f64(u, batch)
    prop = p.p_bpm
    shift = prop isa ShiftProp
    k = prop.optical_components[2]
    ds = kernel_ds(k, u)
    ns = size(p.dn, ndims(p.dn))
    dz = Float64(k.z)
    # kdz was fixed from the template's tilts at construction (bpm.jl:70); the library derives it
    # from the tilts it is given, so the field must carry the template's tilts
    kdz = Float64.(collect(vec(p.kdz isa AbstractArray ? Array(p.kdz) : [p.kdz])))
    want = [2π * dz / (cos(tx[b]) * cos(ty[b])) for b in 1:batch]
    all(isapprox.(length(kdz) == 1 ? fill(kdz[1], batch) : kdz, want; rtol = 1e-6)) ||
        error("libfluxhip: the field's tilts differ from the BPM template's, for which kdz was computed")
    flags = shift ? FH_PROP_SHIFT :
            ((k.is_paraxial ? FH_PROP_PARAXIAL : Cint(0)) | (k.track_tilts ? FH_PROP_TRACK_TILTS : Cint(0)))
    n0 = shift ? 1.0 : Float64(k.n0)
    cache = get!(BPMS, p) do; Dict{Any, BpmPtr}() end
    get!(cache, (AMDGPU.device_id(AMDGPU.device()), batch, lams, tx, ty, ds)) do
        r = Ref{BpmPtr}(C_NULL)
        GC.@preserve lams tx ty begin
            d = Ref(PropDesc(nx, ny, batch, ds[1], ds[2], pointer(lams), length(lams), C_NULL, 0,
                             pointer(tx), pointer(ty), batch, n0, flags))
            check(ccall((:fh_bpm_create_ex, libfluxhip), Cint,
                        (CtxPtr, Ptr{PropDesc}, Cdouble, Cint, Cint, Ptr{BpmPtr}),
                        context(), d, dz * ns, ns, BPM_CHECKPOINT[], r), :fh_bpm_create_ex)
        end
        r[]
    end
end

function bpm_forward!(u::DField, p::BPM, direction, save::Bool)
    check(ccall((:fh_bpm_forward, libfluxhip), Cint, (BpmPtr, Ptr{Cvoid}, Ptr{Cfloat}, Cint, Cint),
                bpm_obj(p, u), pointer(u.electric), pointer(p.dn), dir_sign(direction), Cint(save)), :fh_bpm_forward)
    u
end

function bpm_backward!(u::DField, p::BPM, direction, ddn)
    check(ccall((:fh_bpm_backward, libfluxhip), Cint, (BpmPtr, Ptr{Cvoid}, Ptr{Cfloat}, Ptr{Cfloat}, Cint),
                bpm_obj(p, u), pointer(u.electric), pointer(p.dn), ddn === nothing ? C_NULL : pointer(ddn),
                dir_sign(direction)), :fh_bpm_backward)
    u
end

propagate!(u::DField, p::BPM, direction::Type{<:Direction}; u_saved = nothing) =
    bpm_forward!(u, p, direction, u_saved !== nothing)
propagate_and_save!(u::DField, p::BPM{Trainable{Buffered}}, direction::Type{<:Direction}) =
    bpm_forward!(u, p, direction, true)
propagate_and_save!(u::DField, u_saved::ROCArray, p::BPM{Trainable{Unbuffered}}, direction::Type{<:Direction}) =
    bpm_forward!(u, p, direction, true)
alloc_saved_buffer(u::DField, p::BPM{Trainable{Unbuffered}}) = similar(u.electric, 0)
backpropagate!(u::DField, p::BPM, direction::Type{<:Direction}; u_saved = nothing, ∂p = nothing) =
    bpm_backward!(u, p, direction, ∂p === nothing ? nothing : ∂p.dn)
function backpropagate_with_gradient!(∂v::DField, u_saved::AbstractArray, ∂p::NamedTuple, p::BPM{<:Trainable},
                                      direction::Type{<:Direction})
    bpm_backward!(∂v, p, direction, ∂p.dn)
    (∂v, ∂p)
end

# SquaredIntensityDifference (squared_intensity_difference.jl:78-90) on the device: the values
# (FP64, one per field or one for a spatial-only target) and, in the pullback, 4 dc d u.
function sq_intensity_diff!(loss::ROCArray{Float64}, du::Union{Nothing, ROCArray{ComplexF32}},
                            u::DField, target::ROCArray{Float32},
                            dc::Union{Nothing, ROCArray{Float64}} = nothing)
    nx, ny, batch = grid(u)
    tb = div(length(target), nx * ny)
    check(ccall((:fh_sq_intensity_diff, libfluxhip), Cint,
                (CtxPtr, Ptr{Cvoid}, Ptr{Cfloat}, Cint, Cint, Cint, Cint, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cvoid}),
                context(), pointer(u.electric), pointer(target), nx, ny, batch, tb,
                dc === nothing ? C_NULL : pointer(dc), pointer(loss), du === nothing ? C_NULL : pointer(du)),
          :fh_sq_intensity_diff)
    loss
end

end # module

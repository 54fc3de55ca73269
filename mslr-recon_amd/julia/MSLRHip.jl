#=
MSLRHip.jl
Julia binding of libmslr_hip.so (the C ABI in include/mslr.h) for the
reference's driver: the same call sites as Asense_gpu / patchSVST_gpu /
pogm_mod / poweriter_mod (src/sense_gpu.jl, src/recon.jl, src/mirt_mod.jl,
scripts/reconstruct.jl), with device buffers owned by the library.

Load with:  include("MSLRHip.jl"); using .MSLRHip
Set ENV["MSLR_HIP_LIB"] to the path of libmslr_hip.so if it is not next to
this file's package directory (../mslr_hip/libmslr_hip.so).
=#
module MSLRHip

using LinearMapsAA: LinearMapAA   # a dependency of the reference (Project.toml)

export HipContext, HipArray, Asense_hip, Asense_hip_frames, img2patches_hip, patches2img_hip, SVST_hip,
       patchSVST_hip, patch_nucnorm_hip, pogm_hip, poweriter_hip, upload, download,
       set_products!, SVST_vjp_hip, pogm_checkpointed_hip, lambda_vjp_hip, pogm_lambda_grad_hip,
       tSNR_hip, detrend_hip!, set_tuning!, get_tuning

const LIB = get(ENV, "MSLR_HIP_LIB", joinpath(@__DIR__, "..", "mslr_hip", "libmslr_hip.so"))

# status codes -> the reference's exception kinds
function check(code::Cint)
    code == 0 && return nothing
    msg = unsafe_string(ccall((:mslr_last_error, LIB), Cstring, ()))
    code == 1 && throw(ArgumentError(msg))
    code == 2 && throw(AssertionError(msg))
    code == 4 && throw(OutOfMemoryError())
    code == 6 && throw(ErrorException("numeric failure: " * msg))
    error(msg)   # 3: device, 5: unsupported size
end

# ---------------------------------------------------------------- context + buffers
mutable struct HipContext
    h::Ptr{Cvoid}
    function HipContext(device::Integer = 0)
        r = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:mslr_ctx_create, LIB), Cint, (Cint, Ref{Ptr{Cvoid}}), device, r))
        c = new(r[])
        finalizer(c) do x
            x.h != C_NULL && ccall((:mslr_ctx_destroy, LIB), Cint, (Ptr{Cvoid},), x.h)
            x.h = C_NULL
        end
        return c
    end
end

const _CTX = Ref{Union{Nothing,HipContext}}(nothing)
default_context() = (_CTX[] === nothing && (_CTX[] = HipContext(0)); _CTX[])

"""
    set_products!(ctx, products::Symbol)

SVST Gram / reconstruction products of `SVST_hip` and `patchSVST_hip` on `ctx`:
`:fp32` (FP64 Gram products of the FP32 data, the default) or `:split_bf16` (the
mixed-precision form: hi + lo BF16 on the matrix cores).
"""
function set_products!(ctx::HipContext, products::Symbol)
    code = products === :fp32 ? 0 : products === :split_bf16 ? 1 :
           throw(ArgumentError("products must be :fp32 or :split_bf16"))
    check(ccall((:mslr_ctx_set_products, LIB), Cint, (Ptr{Cvoid}, Cint), ctx.h, code))
    return ctx
end

# mslr_tuning (include/mslr.h): launch shapes
struct Tuning
    gram_chunk::Cint
    eig_lanes_limit::Cint
    svst_apply::Cint
    sense_chunks::Cint
    yz_threads_fwd::Cint
    yz_threads_adj::Cint
    update_lines::Cint
    sb2st_schedule::Cint
    apply_fp32::Cint
    reserved::NTuple{3,Cint}
end

"""
    set_tuning!(ctx; gram_chunk=0, eig_lanes_limit=0, svst_apply=0, sense_chunks=0,
                yz_threads_fwd=0, yz_threads_adj=0, update_lines=0, sb2st_schedule=0,
                apply_fp32=0)

Launch shapes on `ctx` (0 keeps the built-in choice), see `mslr_tuning` in
include/mslr.h.  The Python autotuner (`mslr_hip.autotune.autotune_pogm`) picks them
from timed POGM iterations; its JSON cache holds the values to pass here.
"""
function set_tuning!(ctx::HipContext; gram_chunk::Integer = 0, eig_lanes_limit::Integer = 0,
                     svst_apply::Integer = 0, sense_chunks::Integer = 0, yz_threads_fwd::Integer = 0,
                     yz_threads_adj::Integer = 0, update_lines::Integer = 0, sb2st_schedule::Integer = 0,
                     apply_fp32::Integer = 0)
    t = Ref(Tuning(gram_chunk, eig_lanes_limit, svst_apply, sense_chunks, yz_threads_fwd, yz_threads_adj,
                   update_lines, sb2st_schedule, apply_fp32, ntuple(_ -> Cint(0), 3)))
    check(ccall((:mslr_ctx_set_tuning, LIB), Cint, (Ptr{Cvoid}, Ref{Tuning}), ctx.h, t))
    return ctx
end

function get_tuning(ctx::HipContext)
    t = Ref(Tuning(0, 0, 0, 0, 0, 0, 0, 0, 0, ntuple(_ -> Cint(0), 3)))
    check(ccall((:mslr_ctx_get_tuning, LIB), Cint, (Ptr{Cvoid}, Ref{Tuning}), ctx.h, t))
    v = t[]
    return (gram_chunk = Int(v.gram_chunk), eig_lanes_limit = Int(v.eig_lanes_limit),
            svst_apply = Int(v.svst_apply), sense_chunks = Int(v.sense_chunks),
            yz_threads_fwd = Int(v.yz_threads_fwd), yz_threads_adj = Int(v.yz_threads_adj),
            update_lines = Int(v.update_lines), sb2st_schedule = Int(v.sb2st_schedule),
            apply_fp32 = Int(v.apply_fp32))
end

"Column-major device array owned by the library (ComplexF32 unless stated)."
mutable struct HipArray{T,N}
    ptr::Ptr{Cvoid}
    dims::NTuple{N,Int}
    ctx::HipContext
    function HipArray{T}(dims::NTuple{N,Int}; ctx = default_context()) where {T,N}
        r = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:mslr_malloc, LIB), Cint, (Ptr{Cvoid}, Csize_t, Ref{Ptr{Cvoid}}),
                    ctx.h, max(1, prod(dims)) * sizeof(T), r))
        a = new{T,N}(r[], dims, ctx)
        finalizer(a) do x
            x.ptr != C_NULL && ccall((:mslr_free, LIB), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), x.ctx.h, x.ptr)
            x.ptr = C_NULL
        end
        return a
    end
end
Base.size(a::HipArray) = a.dims
Base.eltype(::HipArray{T}) where {T} = T

function upload(x::Array{T,N}; ctx = default_context()) where {T,N}
    a = HipArray{T}(size(x); ctx)
    check(ccall((:mslr_upload, LIB), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{T}, Csize_t),
                ctx.h, a.ptr, x, sizeof(x)))
    return a
end

function download(a::HipArray{T,N}) where {T,N}
    x = Array{T,N}(undef, a.dims)
    check(ccall((:mslr_download, LIB), Cint, (Ptr{Cvoid}, Ptr{T}, Ptr{Cvoid}, Csize_t),
                a.ctx.h, x, a.ptr, sizeof(x)))
    return x
end
Base.Array(a::HipArray) = download(a)

# ---------------------------------------------------------------- SENSE (src/sense_gpu.jl:56-108)
"""
    Asense_hip(samp, smaps; fft_forward=true, unitary=true)

All-frames SENSE operator for samp::AbstractArray{Bool,4} (Nx,Ny,Nz,Nt) (a 3D
mask gives one frame).  `A * x` and `A' * y` act on HipArrays (or host Arrays)
of shape idim = (Nx,Ny,Nz,Nt) and odim = (K,Nc,Nt).  This replaces
`block_diag([Asense_gpu(s, smaps) for s in eachslice(Ω; dims=4)]...)`
(scripts/reconstruct.jl:142).  The adjoint is the exact conjugate transpose.
"""
mutable struct Asense_hip
    h::Ptr{Cvoid}
    N::NTuple{3,Int}
    Nc::Int
    Nt::Int
    K::Int
    ctx::HipContext
end

function Asense_hip(samp::AbstractArray{Bool}, smaps::AbstractArray{<:Complex,4};
                    fft_forward::Bool = true, unitary::Bool = true, ctx = default_context())
    m = ndims(samp) == 3 ? reshape(samp, size(samp)..., 1) : samp
    N = size(smaps)[1:3]
    Nc = size(smaps, 4)
    Nt = size(m, 4)
    s32 = ComplexF32.(collect(smaps))
    m8 = UInt8.(collect(m))
    r = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:mslr_sense_create, LIB), Cint,
                (Ptr{Cvoid}, Ptr{Int64}, Int64, Int64, Ptr{ComplexF32}, Ptr{UInt8}, Cint, Cint, Ref{Ptr{Cvoid}}),
                ctx.h, Int64[N...], Nc, Nt, s32, m8, fft_forward, unitary, r))
    K = ccall((:mslr_sense_nsamples, LIB), Int64, (Ptr{Cvoid},), r[])
    A = Asense_hip(r[], N, Nc, Nt, K, ctx)
    finalizer(A) do x
        x.h != C_NULL && ccall((:mslr_sense_destroy, LIB), Cint, (Ptr{Cvoid},), x.h)
        x.h = C_NULL
    end
    return A
end

struct AdjointHip
    A::Asense_hip
end
Base.adjoint(A::Asense_hip) = AdjointHip(A)
idim(A::Asense_hip) = (A.N..., A.Nt)
odim(A::Asense_hip) = (A.K, A.Nc, A.Nt)

function Base.:*(A::Asense_hip, x::HipArray{ComplexF32})
    y = HipArray{ComplexF32}((A.K, A.Nc, A.Nt); ctx = A.ctx)
    check(ccall((:mslr_sense_forward, LIB), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int64, Int64),
                A.h, x.ptr, y.ptr, 0, A.Nt))
    return y
end
Base.:*(A::Asense_hip, x::Array) = download(A * upload(ComplexF32.(x); ctx = A.ctx))

function Base.:*(Ah::AdjointHip, y::HipArray{ComplexF32})
    A = Ah.A
    x = HipArray{ComplexF32}((A.N..., A.Nt); ctx = A.ctx)
    check(ccall((:mslr_sense_adjoint, LIB), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int64, Int64),
                A.h, y.ptr, x.ptr, 0, A.Nt))
    return x
end
Base.:*(Ah::AdjointHip, y::Array) = download(Ah * upload(ComplexF32.(y); ctx = Ah.A.ctx))

"""
    Asense_hip_frames(Ω, smaps; fft_forward=true, unitary=true) -> Vector of LinearMapAA

One `LinearMapAA` per frame t with idim = (Nx,Ny,Nz), odim = (K,Nc) and T = ComplexF32,
the interface of `Asense_gpu(samp, smaps)` (src/sense_gpu.jl:56-108), so that the
reference's own assembly works unchanged:

    A = block_diag(Asense_hip_frames(Ω, smaps)...)      # scripts/reconstruct.jl:142
    σ1A = poweriter_mod(undim(A))                         # scripts/reconstruct.jl:154

All frames share one all-frames plan (one copy of the coil maps on the device); frame
t's forward and adjoint are `mslr_sense_forward/adjoint` on frames [t-1, t).  Host
arrays in and out, as LinearMapsAA passes them.  (The fast path is `Asense_hip`, which
applies all frames in one call on device arrays.)
"""
function Asense_hip_frames(Ω::AbstractArray{Bool,4}, smaps::AbstractArray{<:Complex,4};
                           fft_forward::Bool = true, unitary::Bool = true, ctx = default_context())
    A = Asense_hip(Ω, smaps; fft_forward, unitary, ctx)
    N, K, Nc = A.N, A.K, A.Nc
    function frame_map(t::Int)
        fwd = function (x)
            xd = upload(reshape(ComplexF32.(x), N..., 1); ctx = A.ctx)
            yd = HipArray{ComplexF32}((K, Nc, 1); ctx = A.ctx)
            check(ccall((:mslr_sense_forward, LIB), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int64, Int64),
                        A.h, xd.ptr, yd.ptr, t - 1, 1))
            return reshape(download(yd), K, Nc)
        end
        adj = function (y)
            yd = upload(reshape(ComplexF32.(y), K, Nc, 1); ctx = A.ctx)
            xd = HipArray{ComplexF32}((N..., 1); ctx = A.ctx)
            check(ccall((:mslr_sense_adjoint, LIB), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int64, Int64),
                        A.h, yd.ptr, xd.ptr, t - 1, 1))
            return reshape(download(xd), N...)
        end
        return LinearMapAA(fwd, adj, (K * Nc, prod(N)); idim = N, odim = (K, Nc), T = ComplexF32)
    end
    return [frame_map(t) for t in 1:A.Nt]
end

"poweriter_mod(undim(A)) in FP64 on the device (src/mirt_mod.jl:215-238); returns (nothing, σ1)."
function poweriter_hip(A::Asense_hip; niter::Int = 200, tol::Real = 1e-6, chat::Bool = true)
    s = Ref{Cdouble}(0); it = Ref{Cint}(0)
    check(ccall((:mslr_sense_poweriter, LIB), Cint, (Ptr{Cvoid}, Cint, Cdouble, Ref{Cdouble}, Ref{Cint}),
                A.h, niter, tol, s, it))
    chat && it[] < niter && @info "Power iteration converged at iter $(it[])"
    return nothing, s[]
end

# ---------------------------------------------------------------- patches + SVST (src/recon.jl)
function patch_geometry(N, ps, ss)
    pc = zeros(Int64, 3); ns = zeros(Int64, 3); Np = Ref{Int64}(0)
    check(ccall((:mslr_patch_geometry, LIB), Cint,
                (Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ref{Int64}),
                Int64[N...], Int64[ps...], Int64[ss...], pc, ns, Np))
    return Tuple(pc), Tuple(ns), Np[]
end

function img2patches_hip(img::HipArray{ComplexF32,4}, patch_size, stride_size)
    N = img.dims[1:3]; Nt = img.dims[4]
    pc, _, Np = patch_geometry(N, patch_size, stride_size)
    P = HipArray{ComplexF32}((prod(pc), Nt, Np); ctx = img.ctx)
    check(ccall((:mslr_img2patches, LIB), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Int64}, Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Cvoid}),
                img.ctx.h, img.ptr, Int64[N...], Nt, Int64[patch_size...], Int64[stride_size...], P.ptr))
    return P
end

function patches2img_hip(P::HipArray{ComplexF32,3}, patch_size, stride_size, og_size)
    Nt = P.dims[2]
    img = HipArray{ComplexF32}((og_size..., Nt); ctx = P.ctx)
    check(ccall((:mslr_patches2img, LIB), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Int64}, Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Cvoid}),
                P.ctx.h, P.ptr, Int64[og_size...], Nt, Int64[patch_size...], Int64[stride_size...], img.ptr))
    return img
end

"SVST(X, β) (src/recon.jl:183-198) in place on a device matrix; returns X."
function SVST_hip(X::HipArray{ComplexF32,2}, β::Real)
    b = Cdouble[β]
    check(ccall((:mslr_svst_batched, LIB), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Int64, Int64, Int64, Ptr{Cdouble}, Cint),
                X.ctx.h, X.ptr, X.dims[1], X.dims[2], 1, b, 0))
    return X
end

"patchSVST(img, β | λs, patch_size, stride_size) (src/recon.jl:251-269)."
function patchSVST_hip(img::HipArray{ComplexF32,4}, β, patch_size, stride_size)
    if β isa AbstractVector
        P = img2patches_hip(img, patch_size, stride_size)
        b = Cdouble.(β)
        check(ccall((:mslr_svst_batched, LIB), Cint,
                    (Ptr{Cvoid}, Ptr{Cvoid}, Int64, Int64, Int64, Ptr{Cdouble}, Cint),
                    P.ctx.h, P.ptr, P.dims[1], P.dims[2], P.dims[3], b, 1))
        return patches2img_hip(P, patch_size, stride_size, img.dims[1:3])
    end
    out = HipArray{ComplexF32}(img.dims; ctx = img.ctx)
    check(ccall((:mslr_patch_svst, LIB), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Int64}, Int64, Ptr{Int64}, Ptr{Int64}, Cdouble),
                img.ctx.h, img.ptr, out.ptr, Int64[img.dims[1:3]...], img.dims[4],
                Int64[patch_size...], Int64[stride_size...], β))
    return out
end

"""
    SVST_vjp_hip(Z, Ybar, β) -> (Zbar, βbar)

Pullback of `Y = SVST(Z, β)` (src/recon.jl:183-205) for device matrices Z, Ybar of
size (p, Nt) or (p, Nt, Np) (any p and Nt): Zbar and βbar = ∂⟨Ybar, Y⟩/∂β (one per
patch when β is a vector).  The rrule of SVST calls it.
"""
function SVST_vjp_hip(Z::HipArray{ComplexF32}, Ybar::HipArray{ComplexF32}, β)
    Z.dims == Ybar.dims || throw(AssertionError("Z and Ybar differ in shape"))
    p, Nt = Z.dims[1], Z.dims[2]
    Np = length(Z.dims) == 3 ? Z.dims[3] : 1
    per = β isa AbstractVector
    b = per ? Cdouble.(β) : Cdouble[β]
    bb = zeros(Cdouble, per ? Np : 1)
    Zbar = HipArray{ComplexF32}(Z.dims; ctx = Z.ctx)
    check(ccall((:mslr_svst_vjp, LIB), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int64, Int64, Int64, Ptr{Cdouble}, Cint,
                 Ptr{Cdouble}), Z.ctx.h, Z.ptr, Ybar.ptr, Zbar.ptr, p, Nt, Np, b, per, bb))
    return Zbar, per ? bb : bb[1]
end

function patch_nucnorm_hip(P::HipArray{ComplexF32,3})
    r = Ref{Cdouble}(0)
    check(ccall((:mslr_patch_nucnorm, LIB), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Int64, Int64, Int64, Ref{Cdouble}),
                P.ctx.h, P.ptr, P.dims[1], P.dims[2], P.dims[3], r))
    return r[]
end

# patch_nucnorm(P, λs) (src/recon.jl:157-166): one batched pass over all patches
function patch_nucnorm_hip(P::HipArray{ComplexF32,3}, λs::Vector)
    @assert length(λs) == P.dims[3] "need one weight per patch"
    r = Ref{Cdouble}(0)
    lam = Vector{Cdouble}(λs)
    check(ccall((:mslr_patch_nucnorm_weighted, LIB), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Int64, Int64, Int64, Ptr{Cdouble}, Ref{Cdouble}),
                P.ctx.h, P.ptr, P.dims[1], P.dims[2], P.dims[3], lam, r))
    return r[]
end

# ---------------------------------------------------------------- POGM (src/mirt_mod.jl:74-188)
struct PogmDesc
    N::NTuple{3,Int64}
    Nt::Int64
    nscales::Int64
    patch_sizes::Ptr{Int64}
    strides::Ptr{Int64}
    lambdas::Ptr{Cdouble}
    L::Cdouble
    niter::Cint
    mom::Cint
    restart::Cint
    restart_cutoff::Cdouble
    bsig::Cdouble
    form::Cint
    log_costs::Cint
    checkpoint::Cint
end

"""
    pogm_hip(A, ksp, L, λs, PATCH_SIZES, STRIDES; niter, form=0, log_costs=false)
        -> (X, restarts)  or, with log_costs, (X, costs)

The device form of
    pogm_mod(X0, comp_cost, dc_cost_grad, L; mom=:pogm, niter, g_prox, fun)
as called by run_recon (scripts/reconstruct.jl:196-226), X0[..,1] = A'ksp.
`ksp` is (K,Nc,Nt) in the reference's sample order.  form: 0 plain, 1 hipGraph,
2 hipGraph with split-BF16 SVST products (mixed precision).  With log_costs the
second value is the logger's `costs` (niter + 1 tuples (dc_cost, reg_cost,
is_restart), scripts/reconstruct.jl:205,228-230), kept on the device.
"""
function pogm_hip(A::Asense_hip, ksp, L::Real, λs::Vector{Float64}, PATCH_SIZES, STRIDES;
                  niter::Int, form::Integer = 0, log_costs::Bool = false)
    S = length(PATCH_SIZES)
    ps = Int64[p[a] for p in PATCH_SIZES for a in 1:3]
    ss = Int64[s[a] for s in STRIDES for a in 1:3]
    yd = ksp isa HipArray ? ksp : upload(ComplexF32.(ksp); ctx = A.ctx)
    st = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve ps ss λs begin
        d = PogmDesc(Int64.(A.N), A.Nt, S, pointer(ps), pointer(ss), pointer(λs), L, niter, 0, 0, 0.0, 1.0, form,
                     log_costs ? 1 : 0, 0)
        check(ccall((:mslr_pogm_create, LIB), Cint, (Ptr{Cvoid}, Ref{PogmDesc}, Ref{Ptr{Cvoid}}), A.h, d, st))
    end
    try
        check(ccall((:mslr_pogm_init, LIB), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}), st[], yd.ptr, C_NULL))
        check(ccall((:mslr_pogm_run, LIB), Cint, (Ptr{Cvoid}, Cint), st[], niter))
        it = Ref{Cint}(0); rs = zeros(Cint, max(1, niter))
        check(ccall((:mslr_pogm_log, LIB), Cint, (Ptr{Cvoid}, Ref{Cint}, Ptr{Cint}), st[], it, rs))
        xp = ccall((:mslr_pogm_x, LIB), Ptr{Cvoid}, (Ptr{Cvoid},), st[])
        X = Array{ComplexF32}(undef, A.N..., A.Nt, S)
        check(ccall((:mslr_download, LIB), Cint, (Ptr{Cvoid}, Ptr{ComplexF32}, Ptr{Cvoid}, Csize_t),
                    A.ctx.h, X, xp, sizeof(X)))
        restarts = Bool.(rs[1:it[]])
        log_costs || return X, restarts
        n = Ref{Cint}(0); dc = zeros(Cdouble, niter + 1); reg = zeros(Cdouble, niter + 1)
        check(ccall((:mslr_pogm_costs, LIB), Cint, (Ptr{Cvoid}, Ref{Cint}, Ptr{Cdouble}, Ptr{Cdouble}),
                    st[], n, dc, reg))
        flags = [false; restarts]
        return X, [(dc[i], reg[i], flags[i]) for i in 1:n[]]
    finally
        ccall((:mslr_pogm_destroy, LIB), Cint, (Ptr{Cvoid},), st[])
    end
end

"tSNR(img) (src/analysis.jl:39-44) of a device image series; returns the Float32 map."
function tSNR_hip(img::HipArray{ComplexF32})
    V = prod(img.dims[1:end-1])
    out = HipArray{Float32}(img.dims[1:end-1]; ctx = img.ctx)
    check(ccall((:mslr_tsnr, LIB), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Int64, Int64, Ptr{Cvoid}),
                img.ctx.h, img.ptr, V, img.dims[end], out.ptr))
    return out
end

"detrend!(img) (src/analysis.jl:122-132) in place on a device image series."
function detrend_hip!(img::HipArray{ComplexF32})
    V = prod(img.dims[1:end-1])
    check(ccall((:mslr_detrend, LIB), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Int64, Int64),
                img.ctx.h, img.ptr, V, img.dims[end]))
    return img
end

"A POGM state kept on the device with its per-iteration checkpoints (the derivative form)."
mutable struct PogmCheckpointed
    h::Ptr{Cvoid}
    A::Asense_hip
    S::Int
end

"""
    pogm_checkpointed_hip(A, ksp, L, λs, PATCH_SIZES, STRIDES; niter, form=0) -> (X_recon, state)

Runs `niter` POGM iterations as `pogm_hip` does, keeping each iteration's prox
input on the device; `lambda_vjp_hip(state, xbar)` then pulls a cotangent of the
image sum X_recon = Σ_s X_s back to the per-scale weights.
"""
function pogm_checkpointed_hip(A::Asense_hip, ksp, L::Real, λs::Vector{Float64}, PATCH_SIZES, STRIDES;
                               niter::Int, form::Integer = 0)
    S = length(PATCH_SIZES)
    ps = Int64[p[a] for p in PATCH_SIZES for a in 1:3]
    ss = Int64[s[a] for s in STRIDES for a in 1:3]
    yd = ksp isa HipArray ? ksp : upload(ComplexF32.(ksp); ctx = A.ctx)
    r = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve ps ss λs begin
        d = PogmDesc(Int64.(A.N), A.Nt, S, pointer(ps), pointer(ss), pointer(λs), L, niter, 0, 0, 0.0, 1.0, form,
                     0, 1)
        check(ccall((:mslr_pogm_create, LIB), Cint, (Ptr{Cvoid}, Ref{PogmDesc}, Ref{Ptr{Cvoid}}), A.h, d, r))
    end
    st = PogmCheckpointed(r[], A, S)
    finalizer(st) do x
        x.h != C_NULL && ccall((:mslr_pogm_destroy, LIB), Cint, (Ptr{Cvoid},), x.h)
        x.h = C_NULL
    end
    check(ccall((:mslr_pogm_init, LIB), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}), st.h, yd.ptr, C_NULL))
    check(ccall((:mslr_pogm_run, LIB), Cint, (Ptr{Cvoid}, Cint), st.h, niter))
    xp = ccall((:mslr_pogm_x, LIB), Ptr{Cvoid}, (Ptr{Cvoid},), st.h)
    X = Array{ComplexF32}(undef, A.N..., A.Nt, S)
    check(ccall((:mslr_download, LIB), Cint, (Ptr{Cvoid}, Ptr{ComplexF32}, Ptr{Cvoid}, Csize_t),
                A.ctx.h, X, xp, sizeof(X)))
    return dropdims(sum(X; dims = 5); dims = 5), st
end

"λbar[k] = ∂⟨xbar, Σ_s X_s⟩/∂λ_k through the checkpointed iterations (restarts held fixed)."
function lambda_vjp_hip(st::PogmCheckpointed, xbar)
    xb = xbar isa HipArray ? xbar : upload(ComplexF32.(xbar); ctx = st.A.ctx)
    λbar = zeros(Cdouble, st.S)
    check(ccall((:mslr_pogm_lambda_vjp, LIB), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cdouble}),
                st.h, xb.ptr, λbar))
    return λbar
end

"""
    pogm_lambda_grad_hip(A, ksp, L, λs, PATCH_SIZES, STRIDES, x_ref; niter, form=0)
        -> (loss, ∇λ)

MSE of the image sum after `niter` unrolled POGM iterations to `x_ref` and its
gradient with respect to the per-scale weights.
"""
function pogm_lambda_grad_hip(A::Asense_hip, ksp, L::Real, λs::Vector{Float64}, PATCH_SIZES, STRIDES,
                              x_ref; niter::Int, form::Integer = 0)
    Xr, st = pogm_checkpointed_hip(A, ksp, L, λs, PATCH_SIZES, STRIDES; niter = niter, form = form)
    r = ComplexF64.(Xr) .- ComplexF64.(x_ref)
    loss = sum(abs2, r) / length(r)
    return loss, lambda_vjp_hip(st, (2 / length(r)) .* r)
end

end # module MSLRHip

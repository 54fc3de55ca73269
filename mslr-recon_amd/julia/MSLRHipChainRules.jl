# ChainRulesCore rules for the MI355X library (load after MSLRHip.jl, with
# ChainRulesCore in the environment).  The reference has no autodiff; these make
# the per-scale weights λ_k of run_recon's POGM loop (scripts/reconstruct.jl:163-226)
# and SVST on device matrices (src/recon.jl:183-205) differentiable.
module MSLRHipChainRules

using ChainRulesCore
using ..MSLRHip

"""
    mslr_recon_hip(λs, A, ksp, L, PATCH_SIZES, STRIDES; niter, form=0) -> X_recon

The image sum Σ_s X_s after `niter` POGM iterations with weights λs.  Its rrule
keeps the checkpointed state of the one forward run for the pullback.
"""
function mslr_recon_hip(λs::Vector{Float64}, A, ksp, L, PATCH_SIZES, STRIDES; niter::Int, form::Integer = 0)
    Xr, _ = pogm_checkpointed_hip(A, ksp, L, λs, PATCH_SIZES, STRIDES; niter = niter, form = form)
    return Xr
end

function ChainRulesCore.rrule(::typeof(mslr_recon_hip), λs::Vector{Float64}, A, ksp, L, PATCH_SIZES,
                              STRIDES; niter::Int, form::Integer = 0)
    Xr, st = pogm_checkpointed_hip(A, ksp, L, λs, PATCH_SIZES, STRIDES; niter = niter, form = form)
    pullback(X̄) = (NoTangent(), lambda_vjp_hip(st, unthunk(X̄)), NoTangent(), NoTangent(),
                   NoTangent(), NoTangent(), NoTangent())
    return Xr, pullback
end

function ChainRulesCore.rrule(::typeof(MSLRHip.SVST_hip), X::HipArray{ComplexF32,2}, β::Real)
    Z = upload(MSLRHip.download(X); ctx = X.ctx)  # SVST_hip works in place
    Y = MSLRHip.SVST_hip(X, β)
    function pullback(Ȳ)
        Yb = unthunk(Ȳ)
        Zbar, βbar = SVST_vjp_hip(Z, Yb isa HipArray ? Yb : upload(ComplexF32.(Yb); ctx = X.ctx), β)
        return (NoTangent(), Zbar, βbar)
    end
    return Y, pullback
end

end # module MSLRHipChainRules

#= Weak-dependency extension of LBMengine.jl for the hip backend: a
ChainRulesCore rrule for the swimming-speed loss, so Zygote (or any
ChainRules-aware AD) differentiates through LBM runs on the device.

    swimmingLoss(build, B1, B2, F; ticks, kw...)

`build(B1, B2, F)` returns a fresh model (modelInit + addSquirmer!) whose
squirmers take B1[i], B2[i] and whose force-density field is F.  The primal
is a forward run; the pullback's cotangents come from the discrete adjoint
of the step (lbm_adjoint_run), scaled by the incoming cotangent.  Enabled by
Project.toml's

    [weakdeps]
    ChainRulesCore = "d360d2e6-b24c-11e9-a2a3-2a2ae2dbcce4"
    [extensions]
    LBMengineChainRulesCoreExt = "ChainRulesCore"
=#
module LBMengineChainRulesCoreExt

using LBMengine
using ChainRulesCore

LBMengine.swimmingLoss(build, B1, B2, F; ticks, kw...) = LBMengine.swimmingLoss(build(B1, B2, F); ticks, kw...)

function ChainRulesCore.rrule(::typeof(LBMengine.swimmingLoss), build, B1, B2, F; ticks, kw...)
    model = build(B1, B2, F)
    L, g = LBMengine.swimmingLossGradient(model; ticks, kw...)
    function swimmingLossPullback(ΔL)
        Δ = unthunk(ΔL)
        ΔF = g.forceDensity === nothing ? ZeroTangent() : Δ .* g.forceDensity
        return NoTangent(), NoTangent(), Δ .* g.B1, Δ .* g.B2, ΔF
    end
    return L, swimmingLossPullback
end

end

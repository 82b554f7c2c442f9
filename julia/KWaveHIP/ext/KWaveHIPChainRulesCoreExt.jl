# ChainRulesCore rule of the sensor misfit: the cotangent of the sound-speed
# map comes from the checkpointed discrete adjoint (kw_sim_grad_c0), not from
# differentiating the time loop.
module KWaveHIPChainRulesCoreExt

using ChainRulesCore
using KWaveHIP

function ChainRulesCore.rrule(::typeof(KWaveHIP.misfit), c0::AbstractArray, kgrid, medium, source, sensor, p_obs;
                              kwargs...)
    med = KWaveHIP._with_sound_speed(medium, c0)
    loss, g = KWaveHIP.gradient_c0(kgrid, med, source, sensor, p_obs; kwargs...)
    function misfit_pullback(dl)
        return (NoTangent(), unthunk(dl) .* g, NoTangent(), NoTangent(), NoTangent(), NoTangent(), NoTangent())
    end
    return loss, misfit_pullback
end

end # module

# KWave.jl's own types on the GPU time loop.
#
# `KWaveHIP.kspace_first_order(kgrid, medium, source, sensor; kwargs...)` takes
# the reference's objects and keywords unchanged (first_order.jl:92-111,
# 287-306, 505-524), so switching a script to the GPU is
#     using KWave
#     using KWaveHIP: kspace_first_order
# and the keyword `backend` selects the path: :hip (default) runs the GPU
# time loop, :cpu forwards to the reference's KWave.kspace_first_order.
# The result is KWave.jl's own SimulationOutput.
module KWaveHIPKWaveExt

using KWave
using KWaveHIP

function KWaveHIP.kspace_first_order(kgrid::KWave.AbstractKWaveGrid, medium::KWave.KWaveMedium,
                                     source::KWave.KWaveSource, sensor::KWave.KWaveSensor;
                                     backend::Symbol=:hip, kwargs...)
    backend === :cpu && return KWave.kspace_first_order(kgrid, medium, source, sensor; kwargs...)
    backend === :hip || error("backend must be :hip or :cpu")
    return KWaveHIP._gpu_kspace_first_order(kgrid, medium, source, sensor; output_type=KWave.SimulationOutput,
                                            kwargs...)
end

function KWaveHIP.kspace_first_order_as(kgrid::KWave.AbstractKWaveGrid, medium::KWave.KWaveMedium,
                                        source::KWave.KWaveSource, sensor::KWave.KWaveSensor;
                                        backend::Symbol=:hip, kwargs...)
    backend === :cpu && return KWave.kspace_first_order_as(kgrid, medium, source, sensor; kwargs...)
    backend === :hip || error("backend must be :hip or :cpu")
    return KWaveHIP._gpu_kspace_first_order(kgrid, medium, source, sensor; axisymmetric=true,
                                            output_type=KWave.SimulationOutput, kwargs...)
end

end # module

# Mirror types of the reference's data model, with the reference's field and
# keyword names, for running the drop-in where KWave.jl is not installed.
# fill_desc reads only these field names, so KWave.jl's own types (through
# KWaveHIPKWaveExt) take the same path.
#   KWaveGrid1D/2D/3D  grid.jl:21-108 (the fields the solver reads)
#   make_time!         grid.jl:320-352
#   KWaveMedium        medium.jl:18-76
#   KWaveSource        source.jl:24-73, SourceMode types.jl:14
#   KWaveSensor        sensor.jl:31-38
#   SimulationOutput   sensor.jl:76-87

abstract type AbstractKWaveGrid end

struct KWaveGrid1D <: AbstractKWaveGrid
    Nx::Int
    dx::Float64
    dt::Base.RefValue{Float64}
    Nt::Base.RefValue{Int}
    t_array::Vector{Float64}
end
struct KWaveGrid2D <: AbstractKWaveGrid
    Nx::Int
    dx::Float64
    Ny::Int
    dy::Float64
    dt::Base.RefValue{Float64}
    Nt::Base.RefValue{Int}
    t_array::Vector{Float64}
end
struct KWaveGrid3D <: AbstractKWaveGrid
    Nx::Int
    dx::Float64
    Ny::Int
    dy::Float64
    Nz::Int
    dz::Float64
    dt::Base.RefValue{Float64}
    Nt::Base.RefValue{Int}
    t_array::Vector{Float64}
end

KWaveGrid(Nx::Int, dx::Real) = KWaveGrid1D(Nx, Float64(dx), Ref(0.0), Ref(0), Float64[])
KWaveGrid(Nx::Int, dx::Real, Ny::Int, dy::Real) =
    KWaveGrid2D(Nx, Float64(dx), Ny, Float64(dy), Ref(0.0), Ref(0), Float64[])
KWaveGrid(Nx::Int, dx::Real, Ny::Int, dy::Real, Nz::Int, dz::Real) =
    KWaveGrid3D(Nx, Float64(dx), Ny, Float64(dy), Nz, Float64(dz), Ref(0.0), Ref(0), Float64[])

function make_time!(grid::AbstractKWaveGrid, sound_speed; cfl::Float64=0.3, t_end::Union{Nothing, Float64}=nothing)
    c_max, c_min = maximum(sound_speed), minimum(sound_speed)
    ds = _grid_spacing(grid)
    dt = cfl * minimum(ds) / c_max
    grid.dt[] = dt
    if t_end === nothing
        diag = sqrt(sum((n * d)^2 for (n, d) in zip(_grid_size(grid), ds)))
        t_end = 2.0 * diag / c_min
    end
    Nt = ceil(Int, t_end / dt) + 1
    grid.Nt[] = Nt
    resize!(grid.t_array, Nt)
    grid.t_array .= (0:Nt-1) .* dt
    return dt, Nt
end

struct KWaveMedium{T<:AbstractFloat}
    sound_speed::Union{T, AbstractArray{T}}
    density::Union{T, AbstractArray{T}}
    alpha_coeff::Union{Nothing, T, AbstractArray{T}}
    alpha_power::Union{Nothing, T}
    alpha_mode::Symbol
    BonA::Union{Nothing, T, AbstractArray{T}}
end
_f64(x) = x === nothing ? nothing : (x isa Real ? Float64(x) : Float64.(x))
KWaveMedium(; sound_speed, density=1.0, alpha_coeff=nothing, alpha_power=nothing,
            alpha_mode::Symbol=:no_absorption, BonA=nothing) =
    KWaveMedium{Float64}(_f64(sound_speed), _f64(density), _f64(alpha_coeff),
                         alpha_power === nothing ? nothing : Float64(alpha_power), alpha_mode, _f64(BonA))

@enum SourceMode Dirichlet Additive AdditiveNoCorrection

struct KWaveSource{T<:AbstractFloat}
    p0::Union{Nothing, AbstractArray{T}}
    p_mask::Union{Nothing, AbstractArray{Bool}}
    p::Union{Nothing, AbstractArray{T}}
    p_mode::SourceMode
    u_mask::Union{Nothing, AbstractArray{Bool}}
    ux::Union{Nothing, AbstractArray{T}}
    uy::Union{Nothing, AbstractArray{T}}
    uz::Union{Nothing, AbstractArray{T}}
    u_mode::SourceMode
end
KWaveSource(; p0=nothing, p_mask=nothing, p=nothing, p_mode::SourceMode=Additive, u_mask=nothing,
            ux=nothing, uy=nothing, uz=nothing, u_mode::SourceMode=Additive) =
    KWaveSource{Float64}(_f64(p0), p_mask, _f64(p), p_mode, u_mask, _f64(ux), _f64(uy), _f64(uz), u_mode)

Base.@kwdef struct KWaveSensor
    mask::Union{Nothing, AbstractArray{Bool}, AbstractMatrix{Float64}} = nothing
    record::Vector{Symbol} = [:p]
    time_reversal_boundary_data::Union{Nothing, AbstractArray} = nothing
    directivity_angle::Union{Nothing, AbstractArray} = nothing
    directivity_size::Union{Nothing, Float64} = nothing
    frequency_response::Union{Nothing, Tuple{Float64, Float64}} = nothing
end

struct SimulationOutput
    data::Dict{Symbol, AbstractArray}
end
Base.getindex(o::SimulationOutput, key::Symbol) = o.data[key]
Base.haskey(o::SimulationOutput, key::Symbol) = haskey(o.data, key)
Base.keys(o::SimulationOutput) = keys(o.data)
Base.show(io::IO, o::SimulationOutput) =
    print(io, "SimulationOutput with fields: ", join(sort(collect(keys(o.data))), ", "))

"""
    KWaveHIP

Julia host side of the MI355X time loop of `kspace_first_order`
(KWave.jl/src/solver/first_order.jl:92-281, 287-499, 505-728).  It keeps the
reference's call

    kspace_first_order(kgrid, medium, source, sensor; pml_inside, pml_size, pml_alpha,
                       smooth_p0, smooth_c0, smooth_rho0, data_cast, save_to_disk,
                       plot_sim, plot_layout, plot_scale, record_movie,
                       progress_callback, show_progress) -> SimulationOutput

and reaches `libkwave_hip.so` through `ccall` on the C ABI of
`include/kwave_hip.h`.  Everything the reference derives before its loop --
validation, `findall` order, `c_ref = mean(c0)`, the alpha-mode mapping,
p0 -- is decided here and written into one `KwSimDesc`; the loop, the
tables and the sensor reductions run in the library.

The objects are read by their reference field names only, so the same code
serves KWave.jl's own `KWaveGrid*`/`KWaveMedium`/`KWaveSource`/`KWaveSensor`
(through the `KWaveHIPKWaveExt` extension) and the mirror types below, which
let the package run where KWave.jl is not installed.
"""
module KWaveHIP

using Libdl

export kspace_first_order, kspace_first_order_as, kspace_first_order_batch, gradient_c0, misfit, SimulationOutput,
       autotune!, tune_table, tune_load!, tune_clear!,
       KWaveGrid, KWaveGrid1D, KWaveGrid2D, KWaveGrid3D, KWaveMedium, KWaveSource, KWaveSensor,
       SourceMode, Dirichlet, Additive, AdditiveNoCorrection, make_time!

# ----------------------------------------------------------------------------
# Library
# ----------------------------------------------------------------------------
const _DEFAULT_LIB = normpath(joinpath(@__DIR__, "..", "..", "..", "k-wave-julia_amd", "kwave_amd",
                                       "libkwave_hip.so"))
libpath() = get(ENV, "KWAVE_HIP_LIB", _DEFAULT_LIB)
const _lib = Ref{Ptr{Cvoid}}(C_NULL)

function lib()
    if _lib[] == C_NULL
        isfile(libpath()) || error("libkwave_hip.so not found at $(libpath()); build it with " *
                                   "`python k-wave-julia_amd/csrc/build.py` (hipcc, gfx950) or set KWAVE_HIP_LIB")
        _lib[] = Libdl.dlopen(libpath())
        v = ccall(Libdl.dlsym(_lib[], :kw_abi_version), Int32, ())
        v == 1 || error("libkwave_hip ABI version mismatch ($v)")
        # a library built from another version of kwave_hip.h would misread the descriptor
        ns, ne = Ref{Int64}(0), Ref{Int64}(0)
        ccall(Libdl.dlsym(_lib[], :kw_desc_sizes), Int32, (Ref{Int64}, Ref{Int64}), ns, ne)
        ns[] == sizeof(KwSimDesc) || error("$(libpath()) was built from a different kwave_hip.h " *
                                           "(kw_sim_desc is $(ns[]) bytes there, $(sizeof(KwSimDesc)) here): rebuild it")
    end
    return _lib[]
end
sym(name::Symbol) = Libdl.dlsym(lib(), name)

const KW_OK = Int32(0)

struct KWaveHIPError <: Exception
    code::Int32
    msg::String
end
Base.showerror(io::IO, e::KWaveHIPError) = print(io, "KWaveHIPError(", e.code, "): ", e.msg)

function check(code::Integer)
    if code != KW_OK
        msg = unsafe_string(ccall(sym(:kw_last_error), Cstring, ()))
        throw(KWaveHIPError(Int32(code), isempty(msg) ? "libkwave_hip error $code" : msg))
    end
    return nothing
end

# Record fields, kw_record bits (sensor.jl:41-45)
const REC = Dict{Symbol, UInt32}(:p => 1 << 0, :p_max => 1 << 1, :p_min => 1 << 2, :p_rms => 1 << 3,
                                 :p_final => 1 << 4, :ux => 1 << 5, :uy => 1 << 6, :uz => 1 << 7,
                                 :u_max => 1 << 8, :u_rms => 1 << 9, :u_final => 1 << 10,
                                 :I_avg => 1 << 11, :I_max => 1 << 12)
const VALID_RECORD_FIELDS = Set(keys(REC))
const FIELD = Dict{Symbol, Int32}(:p => 0, :ux => 1, :uy => 2, :uz => 3, :rhox => 4, :rhoy => 5, :rhoz => 6,
                                   :rho_prev => 7)
# alpha_mode symbols (medium.jl:66); any other symbol acts as :stokes (first_order_steps.jl:252)
alpha_mode_code(m::Symbol) = m === :no_absorption ? Int32(0) : (m === :no_dispersion ? Int32(1) : Int32(2))

# ----------------------------------------------------------------------------
# kw_sim_desc (include/kwave_hip.h), field for field and in order
# ----------------------------------------------------------------------------
Base.@kwdef struct KwSimDesc
    ndim::Int32 = 0
    N::NTuple{3, Int64} = (1, 1, 1)
    d::NTuple{3, Float64} = (1.0, 1.0, 1.0)
    dt::Float64 = 0.0
    Nt::Int64 = 0
    precision::Int32 = 8
    map_storage::Int32 = 0
    pml_size::NTuple{3, Int32} = (0, 0, 0)
    pml_alpha::NTuple{3, Float64} = (0.0, 0.0, 0.0)
    c_ref::Float64 = 0.0
    c0_offset::Float64 = 0.0
    rho0_offset::Float64 = 0.0
    c0::Ptr{Float64} = C_NULL
    c0_scalar::Float64 = 0.0
    rho0::Ptr{Float64} = C_NULL
    rho0_scalar::Float64 = 0.0
    nonlinear::Int32 = 0
    bona::Ptr{Float64} = C_NULL
    bona_scalar::Float64 = 0.0
    absorbing::Int32 = 0
    alpha_mode::Int32 = 0
    alpha_coeff::Float64 = 0.0
    alpha_power::Float64 = 0.0
    p0::Ptr{Float64} = C_NULL
    smooth_p0::Int32 = 0
    n_psrc::Int64 = 0
    psrc_index::Ptr{Int64} = C_NULL
    psrc_signal::Ptr{Float64} = C_NULL
    psrc_ncols::Int64 = 0
    p_mode::Int32 = 1
    n_usrc::Int64 = 0
    usrc_index::Ptr{Int64} = C_NULL
    usrc_signal::NTuple{3, Ptr{Float64}} = (C_NULL, C_NULL, C_NULL)
    usrc_ncols::NTuple{3, Int64} = (0, 0, 0)
    u_mode::Int32 = 1
    n_sensor::Int64 = 0
    sensor_index::Ptr{Int64} = C_NULL
    record::UInt32 = 0
    tr_data::Ptr{Float64} = C_NULL
    directivity::Int32 = 0
    directivity_angle::Ptr{Float64} = C_NULL
    directivity_angle_scalar::Float64 = 0.0
    directivity_size::Float64 = 0.0
    freq_response::Int32 = 0
    freq_center::Float64 = 0.0
    freq_bandwidth::Float64 = 0.0
    axisymmetric::Int32 = 0
end

include("types.jl")

# ----------------------------------------------------------------------------
# Descriptor fill: what first_order.jl:311-440 / 529-669 derive before the loop
# ----------------------------------------------------------------------------
_grid_size(g) = hasproperty(g, :Nz) ? (g.Nx, g.Ny, g.Nz) : (hasproperty(g, :Ny) ? (g.Nx, g.Ny) : (g.Nx,))
_grid_spacing(g) = hasproperty(g, :dz) ? (g.dx, g.dy, g.dz) : (hasproperty(g, :dy) ? (g.dx, g.dy) : (g.dx,))
_findall0(mask) = Int64.(findall(vec(mask)) .- 1)        # findall order (column-major), 0-based
_colmajor(a) = Vector{Float64}(vec(Float64.(a)))
_pad3(t, fill) = ntuple(i -> i <= length(t) ? t[i] : fill, 3)
_tuple(v, nd) = v isa Tuple ? (length(v) == nd ? v : error("expected a $nd-tuple, got $v")) : ntuple(_ -> v, nd)

"""
FP16 offset storage of a static map: (dequantised Float64 map, offset); the
device stores Float16(a - offset) rounded to nearest and reconstructs
offset + Float16 in the kernel precision.
"""
function quantize_map(a)
    off = round((minimum(a) + maximum(a)) / 2)
    return Float64.(Float16.(a .- off)) .+ off, off
end

"""
    fill_desc(kgrid, medium, source, sensor; kwargs...) -> (desc, keep, info)

`keep` holds every array `desc` points to; wrap the library call that reads
`desc` in `GC.@preserve keep`.
"""
function fill_desc(kgrid, medium, source, sensor; pml_size=20, pml_alpha=2.0, smooth_p0::Bool=true,
                   data_cast::Type=Float64, medium_storage=nothing, axisymmetric::Bool=false)
    if kgrid.Nt[] == 0 || kgrid.dt[] == 0                              # first_order.jl:311-313
        error("Time array not set. Call make_time!(kgrid, sound_speed) before running.")
    end
    for r in sensor.record                                              # sensor.jl:52-58
        r in VALID_RECORD_FIELDS || error("Invalid sensor record field: :$r. Valid fields: " *
                                          join(sort(collect(VALID_RECORD_FIELDS)), ", "))
    end
    shape = _grid_size(kgrid)
    nd = length(shape)
    axisymmetric && nd != 2 && error("kspace_first_order_as takes a 2-D grid (x axial, y radial)")
    pml_alpha isa Integer && throw(MethodError(kspace_first_order, (pml_alpha,)))   # first_order.jl:294
    keep = Any[]
    f = Dict{Symbol, Any}()
    f[:ndim] = Int32(nd)
    f[:N] = _pad3(Int64.(shape), Int64(1))
    f[:d] = _pad3(Float64.(_grid_spacing(kgrid)), 1.0)
    f[:dt] = Float64(kgrid.dt[])
    Nt = Int64(kgrid.Nt[])
    f[:precision] = data_cast === Float32 ? Int32(4) : (data_cast === Float64 ? Int32(8) :
                                                        error("data_cast must be Float32 or Float64"))
    f16 = medium_storage === Float16
    medium_storage === nothing || f16 || error("medium_storage must be Float16 (or nothing)")
    f[:map_storage] = f16 ? Int32(2) : Int32(0)
    f[:pml_size] = _pad3(Int32.(_tuple(pml_size, nd)), Int32(0))
    f[:pml_alpha] = _pad3(Float64.(_tuple(pml_alpha, nd)), 0.0)

    c0, rho0 = medium.sound_speed, medium.density
    if c0 isa Real
        f[:c0_scalar] = Float64(c0)
        f[:c_ref] = Float64(c0)
    else
        size(c0) == shape || error("sound_speed array must match the grid size")
        if f16
            c0, f[:c0_offset] = quantize_map(c0)
        end
        cm = _colmajor(c0)
        push!(keep, cm)
        f[:c0] = pointer(cm)
        f[:c_ref] = sum(c0) / length(c0)                              # mean(c0), first_order.jl:361
    end
    if rho0 isa Real
        f[:rho0_scalar] = Float64(rho0)
    else
        size(rho0) == shape || error("density array must match the grid size")
        if f16
            rho0, f[:rho0_offset] = quantize_map(rho0)
        end
        rm = _colmajor(rho0)
        push!(keep, rm)
        f[:rho0] = pointer(rm)
    end
    if medium.BonA !== nothing
        f[:nonlinear] = Int32(1)
        if medium.BonA isa Real
            f[:bona_scalar] = Float64(medium.BonA)
        else
            bm = _colmajor(medium.BonA .* ones(shape))                # an array B/A broadcasts
            push!(keep, bm)
            f[:bona] = pointer(bm)
        end
    end
    lossless = medium.alpha_coeff === nothing || medium.alpha_mode === :no_absorption   # medium.jl:90
    if !lossless
        medium.alpha_power === nothing && error("alpha_power must be set when alpha_coeff is set")
        medium.alpha_coeff isa Real || error("array alpha_coeff is not supported (first_order.jl:588-592)")
        f[:absorbing] = Int32(1)
        f[:alpha_mode] = alpha_mode_code(medium.alpha_mode)
        f[:alpha_coeff] = Float64(medium.alpha_coeff)
        f[:alpha_power] = Float64(medium.alpha_power)
    end
    f[:axisymmetric] = Int32(axisymmetric)
    trd = sensor.time_reversal_boundary_data
    tr = trd !== nothing && !axisymmetric
    if tr
        Nt = Int64(size(trd, 2))                                        # first_order.jl:430-431
    end
    f[:Nt] = Nt

    if source.p0 !== nothing && !tr
        size(source.p0) == shape || error("p0 must match the grid size")
        p0 = _colmajor(source.p0)
        push!(keep, p0)
        f[:p0] = pointer(p0)
        if smooth_p0                                                    # first_order_steps.jl:70,109,157
            f[:smooth_p0] = Int32(nd == 1 ? shape[1] > 4 : all(>(1), shape))
        end
    end
    if source.p_mask !== nothing && source.p !== nothing
        idx = _findall0(source.p_mask)
        sig = source.p isa AbstractVector ? reshape(source.p, :, 1) : source.p
        size(sig, 1) < length(idx) && throw(BoundsError(sig, (length(idx), 1)))
        sig = Matrix{Float64}(sig[1:length(idx), :])
        push!(keep, idx, sig)
        f[:n_psrc] = length(idx)
        f[:psrc_index] = pointer(idx)
        f[:psrc_signal] = pointer(sig)
        f[:psrc_ncols] = size(sig, 2)
        f[:p_mode] = Int32(source.p_mode)
    end
    if source.u_mask !== nothing
        idx = _findall0(source.u_mask)
        push!(keep, idx)
        f[:n_usrc] = length(idx)
        f[:usrc_index] = pointer(idx)
        ptrs = Ptr{Float64}[C_NULL, C_NULL, C_NULL]
        cols = Int64[0, 0, 0]
        for (a, s) in enumerate((source.ux, source.uy, source.uz)[1:nd])
            s === nothing && continue
            s = s isa AbstractVector ? reshape(s, :, 1) : s
            size(s, 1) < length(idx) && throw(BoundsError(s, (length(idx), 1)))
            s = Matrix{Float64}(s[1:length(idx), :])
            push!(keep, s)
            ptrs[a] = pointer(s)
            cols[a] = size(s, 2)
        end
        f[:usrc_signal] = Tuple(ptrs)
        f[:usrc_ncols] = Tuple(cols)
        f[:u_mode] = Int32(source.u_mode)
    end

    info = Dict{Symbol, Any}(:n_sensor => 0, :cartesian => false, :tr => tr, :Nt => Nt, :shape => shape)
    rec = UInt32(0)
    for r in sensor.record
        rec |= REC[r]
    end
    f[:record] = rec
    if sensor.mask !== nothing
        m = sensor.mask
        if eltype(m) == Bool
            idx = _findall0(m)
            push!(keep, idx)
            f[:n_sensor] = length(idx)
            f[:sensor_index] = isempty(idx) ? Ptr{Int64}(C_NULL) : pointer(idx)
            info[:n_sensor] = length(idx)
        else
            f[:n_sensor] = size(m, 2)                                  # Cartesian: never filled (A-13)
            info[:n_sensor] = size(m, 2)
            info[:cartesian] = true
        end
    end
    if tr
        size(trd, 1) < info[:n_sensor] && throw(BoundsError(trd, (info[:n_sensor], 1)))   # first_order.jl:455
        trc = Matrix{Float64}(trd)
        push!(keep, trc)
        f[:tr_data] = pointer(trc)
    end
    if sensor.directivity_angle !== nothing && sensor.directivity_size !== nothing
        f[:directivity] = Int32(1)
        ang = sensor.directivity_angle
        if ang isa Real || length(ang) == 1
            f[:directivity_angle_scalar] = Float64(first(ang))
        else
            a = _colmajor(ang)
            length(a) < info[:n_sensor] && throw(BoundsError(a, info[:n_sensor]))
            push!(keep, a)
            f[:directivity_angle] = pointer(a)
        end
        f[:directivity_size] = Float64(sensor.directivity_size)
    end
    if sensor.frequency_response !== nothing
        f[:freq_response] = Int32(1)
        f[:freq_center], f[:freq_bandwidth] = Float64.(sensor.frequency_response)
    end
    return KwSimDesc(; f...), keep, info
end

# ----------------------------------------------------------------------------
# Device context and simulation handle
# ----------------------------------------------------------------------------
const _ctx = Dict{Int, Ptr{Cvoid}}()
function context(device::Integer=0)
    get!(_ctx, Int(device)) do
        h = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall(sym(:kw_ctx_create), Int32, (Int32, Ref{Ptr{Cvoid}}), device, h))
        h[]
    end
end

mutable struct Simulation
    h::Ptr{Cvoid}
    desc::KwSimDesc
end

function Simulation(desc::KwSimDesc, keep; device::Integer=0)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve keep begin
        check(ccall(sym(:kw_sim_create), Int32, (Ptr{Cvoid}, Ref{KwSimDesc}, Ref{Ptr{Cvoid}}),
                    context(device), desc, h))
    end
    s = Simulation(h[], desc)
    finalizer(close!, s)
    return s
end

# Stacked firings (kw_sim_create_firings): 2-D descriptors sharing grid and medium, run as one simulation
function Simulation(descs::Vector{KwSimDesc}, keeps; device::Integer=0)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve descs keeps begin
        check(ccall(sym(:kw_sim_create_firings), Int32, (Ptr{Cvoid}, Ptr{KwSimDesc}, Int32, Ref{Ptr{Cvoid}}),
                    context(device), descs, Int32(length(descs)), h))
    end
    s = Simulation(h[], descs[1])
    finalizer(close!, s)
    return s
end

run!(s::Simulation, n::Integer) = check(ccall(sym(:kw_sim_run), Int32, (Ptr{Cvoid}, Int64), s.h, n))
synchronize(s::Simulation) = check(ccall(sym(:kw_sim_synchronize), Int32, (Ptr{Cvoid},), s.h))
function close!(s::Simulation)
    if s.h != C_NULL
        ccall(sym(:kw_sim_destroy), Int32, (Ptr{Cvoid},), s.h)
        s.h = C_NULL
    end
    return nothing
end

_shape(s::Simulation) = Tuple(Int.(s.desc.N[1:s.desc.ndim]))

function field(s::Simulation, name::Symbol)
    out = Vector{Float64}(undef, prod(_shape(s)))
    check(ccall(sym(:kw_sim_get_field), Int32, (Ptr{Cvoid}, Int32, Ptr{Float64}), s.h, FIELD[name], out))
    return reshape(out, _shape(s))
end

# overwrite one field of the state (kw_sim_set_field); with run!(s, 1): one time step from that state
function set_field!(s::Simulation, name::Symbol, a::AbstractArray)
    size(a) == _shape(s) || error("field must have size $(_shape(s))")
    v = Vector{Float64}(vec(a))
    check(ccall(sym(:kw_sim_set_field), Int32, (Ptr{Cvoid}, Int32, Ptr{Float64}), s.h, FIELD[name], v))
    return nothing
end

function sensor_data(s::Simulation, name::Symbol)
    r, c = Ref{Int64}(0), Ref{Int64}(0)
    check(ccall(sym(:kw_sim_sensor_shape), Int32, (Ptr{Cvoid}, UInt32, Ref{Int64}, Ref{Int64}),
                s.h, REC[name], r, c))
    out = Vector{Float64}(undef, r[] * c[])
    check(ccall(sym(:kw_sim_get_sensor), Int32, (Ptr{Cvoid}, UInt32, Ptr{Float64}), s.h, REC[name], out))
    return c[] == 1 ? out : reshape(out, Int(r[]), Int(c[]))
end

# ----------------------------------------------------------------------------
# The drop-in
# ----------------------------------------------------------------------------
"""
    kspace_first_order(kgrid, medium, source, sensor; kwargs...) -> SimulationOutput

GPU run of the reference's `kspace_first_order` (first_order.jl:92/287/505):
same keywords and defaults; `pml_inside`, `smooth_c0`, `smooth_rho0`,
`plot_*`, `record_movie` and `show_progress` are accepted and unused, as in
the reference.  Additional keywords: `medium_storage=Float16` (FP16 maps with
arithmetic in `data_cast`) and `device`.  `save_to_disk` writes the HDF5 input
file the reference writes, through the Python `kwave_amd.hdf5` writer's layout
(not available from Julia here: it raises).  `output_type` picks the
`SimulationOutput` type to return (KWave.jl's own when called through the
extension).
"""
function kspace_first_order(kgrid::AbstractKWaveGrid, medium, source, sensor; kwargs...)
    return _gpu_kspace_first_order(kgrid, medium, source, sensor; output_type=SimulationOutput, kwargs...)
end

function _gpu_kspace_first_order(kgrid, medium, source, sensor; pml_inside::Bool=true, pml_size=20,
                                 pml_alpha=2.0, smooth_p0::Bool=true, smooth_c0::Bool=false,
                                 smooth_rho0::Bool=false, data_cast::Type=Float64, save_to_disk=nothing,
                                 plot_sim::Bool=false, plot_layout::Symbol=:default, plot_scale=:auto,
                                 record_movie=nothing, progress_callback=nothing, show_progress::Bool=true,
                                 medium_storage=nothing, device::Integer=0, axisymmetric::Bool=false,
                                 output_type=SimulationOutput)
    save_to_disk === nothing || error("save_to_disk is written by the Python host (kwave_amd.hdf5)")
    desc, keep, info = fill_desc(kgrid, medium, source, sensor; pml_size=pml_size, pml_alpha=pml_alpha,
                                 smooth_p0=smooth_p0, data_cast=data_cast, medium_storage=medium_storage,
                                 axisymmetric=axisymmetric)
    Nt = info[:Nt]
    info[:tr] && (kgrid.Nt[] = Nt)                                       # first_order.jl:430-431
    sim = Simulation(desc, keep; device=device)
    try
        if progress_callback === nothing
            run!(sim, Nt)
        else
            for t in 1:Nt
                run!(sim, 1)
                progress_callback(t, Nt, field(sim, :p))
            end
        end
        synchronize(sim)
        data = Dict{Symbol, AbstractArray}()
        if sensor.mask !== nothing
            for r in sensor.record
                if r in (:ux, :uy, :uz) && findfirst(==(r), (:ux, :uy, :uz)) > length(info[:shape])
                    data[r] = zeros(Float64, info[:n_sensor], Nt)
                    continue
                end
                data[r] = sensor_data(sim, r)
            end
        end
        if info[:tr]
            data[:p_final] = :p_final in sensor.record ? sensor_data(sim, :p_final) : vec(field(sim, :p))
        end
        return output_type(data)
    finally
        close!(sim)
    end
end

"""
    kspace_first_order_as(kgrid, medium, source, sensor; kwargs...)

Axisymmetric solver (axisymmetric.jl:26-227) on the GPU: 2-D grid, x axial, y radial.
"""
kspace_first_order_as(kgrid::AbstractKWaveGrid, medium, source, sensor; kwargs...) =
    _gpu_kspace_first_order(kgrid, medium, source, sensor; axisymmetric=true, kwargs...)

"""
    gradient_c0(kgrid, medium, sources, sensors, p_obs; checkpoint_every=50, data_cast=Float32, kwargs...)
        -> (loss, dL/dc0)

Loss 1/2 sum (p - p_obs)^2 over the recorded p of every firing and its
gradient with respect to `medium.sound_speed` by the checkpointed discrete
adjoint (`kw_sim_grad_c0`); `sources`, `sensors`, `p_obs` may be vectors (one
entry per firing), whose losses and gradients are summed.  2-D firings run
stacked (`kw_sim_create_firings`, `stack` firings per simulation, default all;
`stack=1` one at a time).
"""
function gradient_c0(kgrid, medium, source, sensor, p_obs; checkpoint_every::Integer=50,
                     data_cast::Type=Float32, device::Integer=0, stack::Union{Nothing, Integer}=nothing, kwargs...)
    medium.sound_speed isa AbstractArray || error("gradient_c0 needs medium.sound_speed as a map")
    firings = source isa AbstractVector ?
              collect(zip(source, sensor isa AbstractVector ? sensor : fill(sensor, length(source)), p_obs)) :
              [(source, sensor, p_obs)]
    total = 0.0
    grad = zeros(Float64, size(medium.sound_speed))
    prepared = map(firings) do (src, sen, obs)
        :p in sen.record || error("the misfit needs :p in sensor.record")
        desc, keep, info = fill_desc(kgrid, medium, src, sen; data_cast=data_cast, kwargs...)
        size(obs) == (info[:n_sensor], info[:Nt]) || error("p_obs must be $((info[:n_sensor], info[:Nt]))")
        (desc, keep, Matrix{Float64}(obs), src)
    end
    # 2-D firings without p0 run stacked: one simulation whose passes cover `stack` firings at once
    stacked = ndims(medium.sound_speed) == 2 && length(prepared) > 1 && stack != 1 &&
              all(x -> x[4].p0 === nothing, prepared)
    groups = stacked ? Iterators.partition(prepared, stack === nothing ? length(prepared) : stack) :
             ([x] for x in prepared)
    for grp in groups
        grp = collect(grp)
        o = reduce(vcat, [x[3] for x in grp])
        g = Vector{Float64}(undef, length(grad))
        loss = Ref{Float64}(0.0)
        sim = length(grp) == 1 && !stacked ? Simulation(grp[1][1], grp[1][2]; device=device) :
              Simulation([x[1] for x in grp], [x[2] for x in grp]; device=device)
        try
            check(ccall(sym(:kw_sim_grad_c0), Int32, (Ptr{Cvoid}, Ptr{Float64}, Int64, Ptr{Float64}, Ref{Float64}),
                        sim.h, o, checkpoint_every, g, loss))
        finally
            close!(sim)
        end
        total += loss[]
        grad .+= reshape(g, size(grad))
    end
    return total, grad
end

# ----------------------------------------------------------------------------
# Batches of independent simulations (kw_batch_*) and the launch-shape tuner
# ----------------------------------------------------------------------------
const KW_BATCH_GRAPH = Int32(1)
const KW_BATCH_PERSISTENT = Int32(2)

# keywords of kspace_first_order: those fill_desc takes, and those the
# reference accepts without using them in the loop (dropped here as there)
const _DESC_KEYS = (:pml_size, :pml_alpha, :smooth_p0, :medium_storage, :axisymmetric)
const _UNUSED_KEYS = (:pml_inside, :smooth_c0, :smooth_rho0, :plot_sim, :plot_layout, :plot_scale, :record_movie,
                      :show_progress)
function _desc_kwargs(kwargs)
    for k in keys(kwargs)
        k in _DESC_KEYS || k in _UNUSED_KEYS || error("unsupported keyword argument $k")
    end
    return (; (k => v for (k, v) in pairs(kwargs) if k in _DESC_KEYS)...)
end

"""
    kspace_first_order_batch(sims; data_cast=Float32, use_graph=true, persistent=false,
                             device=0, kwargs...) -> Vector{SimulationOutput}

Runs a vector of `(kgrid, medium, source, sensor)` tuples as one batch: every
pass of the time step is one launch per line length over all simulations and,
with `use_graph`, one step is one hipGraph replay (`kw_batch_create`);
`persistent` selects the persistent-grid launches.  Keyword arguments are
those of `kspace_first_order` and apply to every simulation; all share
dimensionality and Nt; time reversal is not batched.
"""
function kspace_first_order_batch(sims::AbstractVector; data_cast::Type=Float32, use_graph::Bool=true,
                                  persistent::Bool=false, device::Integer=0, output_type=SimulationOutput,
                                  kwargs...)
    prepared = map(sims) do (kgrid, medium, source, sensor)
        desc, keep, info = fill_desc(kgrid, medium, source, sensor; data_cast=data_cast, _desc_kwargs(kwargs)...)
        info[:tr] && error("time reversal is not batched")
        (desc, keep, info, sensor)
    end
    descs = [x[1] for x in prepared]
    flags = (use_graph ? KW_BATCH_GRAPH : Int32(0)) | (persistent ? KW_BATCH_PERSISTENT : Int32(0))
    h = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve descs prepared begin
        check(ccall(sym(:kw_batch_create), Int32, (Ptr{Cvoid}, Ptr{KwSimDesc}, Int32, Int32, Ref{Ptr{Cvoid}}),
                    context(device), descs, Int32(length(descs)), flags, h))
    end
    outs = Any[]
    try
        Nt = prepared[1][3][:Nt]
        check(ccall(sym(:kw_batch_run), Int32, (Ptr{Cvoid}, Int64), h[], Nt))
        check(ccall(sym(:kw_batch_synchronize), Int32, (Ptr{Cvoid},), h[]))
        for (i, (_, _, info, sensor)) in enumerate(prepared)
            data = Dict{Symbol, AbstractArray}()
            if sensor.mask !== nothing
                for r in sensor.record
                    if r in (:ux, :uy, :uz) && findfirst(==(r), (:ux, :uy, :uz)) > length(info[:shape])
                        data[r] = zeros(Float64, info[:n_sensor], Nt)
                        continue
                    end
                    rows, cols = Ref{Int64}(0), Ref{Int64}(0)
                    check(ccall(sym(:kw_batch_sensor_shape), Int32,
                                (Ptr{Cvoid}, Int32, UInt32, Ref{Int64}, Ref{Int64}), h[], i - 1, REC[r], rows, cols))
                    v = Vector{Float64}(undef, rows[] * cols[])
                    check(ccall(sym(:kw_batch_get_sensor), Int32, (Ptr{Cvoid}, Int32, UInt32, Ptr{Float64}),
                                h[], i - 1, REC[r], v))
                    data[r] = cols[] == 1 ? v : reshape(v, Int(rows[]), Int(cols[]))
                end
            end
            push!(outs, output_type(data))
        end
    finally
        ccall(sym(:kw_batch_destroy), Int32, (Ptr{Cvoid},), h[])
    end
    return outs
end

"""
    autotune!(kgrid, medium, source, sensor; steps=10, data_cast=Float32, device=0, kwargs...) -> String

Times the launch shapes of this simulation's passes (lines per workgroup and,
at line lengths 256 and 512, the plan family) with `kw_autotune` and keeps the
fastest in the library's process-wide table, which later simulations read: the
role FFTW.MEASURE plays in `create_fft_plans` (fft_utils.jl:44-62).  Returns
the table as text (`tune_table`).
"""
function autotune!(kgrid, medium, source, sensor; steps::Integer=10, data_cast::Type=Float32,
                   device::Integer=0, kwargs...)
    desc, keep, _ = fill_desc(kgrid, medium, source, sensor; data_cast=data_cast, _desc_kwargs(kwargs)...)
    GC.@preserve keep begin
        check(ccall(sym(:kw_autotune), Int32, (Ptr{Cvoid}, Ref{KwSimDesc}, Int32, Int32),
                    context(device), desc, Int32(1), Int32(steps)))
    end
    return tune_table()
end

"""
    tune_table() -> String

The tuning table as text, one entry per line: precision bytes, pass class,
line length, lines per workgroup and, where not the default, the plan family.
"""
function tune_table()
    n = Ref{Int64}(0)
    check(ccall(sym(:kw_tune_export), Int32, (Ptr{UInt8}, Int64, Ref{Int64}), C_NULL, 0, n))
    buf = Vector{UInt8}(undef, n[] + 1)
    check(ccall(sym(:kw_tune_export), Int32, (Ptr{UInt8}, Int64, Ref{Int64}), buf, length(buf), n))
    return String(buf[1:n[]])
end

"""Merges a table in `tune_table`'s text form (e.g. read from a cache file)."""
tune_load!(text::AbstractString) = check(ccall(sym(:kw_tune_import), Int32, (Cstring,), text))

"""Empties the tuning table (defaults everywhere)."""
tune_clear!() = check(ccall(sym(:kw_tune_clear), Int32, ()))

# medium with its sound speed replaced (same struct, same other fields)
_with_sound_speed(m, c0) = typeof(m)(Float64.(c0), m.density, m.alpha_coeff, m.alpha_power, m.alpha_mode, m.BonA)

"""
    misfit(c0, kgrid, medium, source, sensor, p_obs; kwargs...) -> loss

1/2 sum (p - p_obs)^2 of the recorded p with `medium`'s sound speed replaced by
`c0` (summed over firings when `source`, `sensor`, `p_obs` are vectors).  With
ChainRulesCore loaded, its `rrule` returns the checkpointed-adjoint gradient
(`gradient_c0`) as the cotangent of `c0`.
"""
function misfit(c0::AbstractArray, kgrid, medium, source, sensor, p_obs; checkpoint_every::Integer=50, kwargs...)
    med = _with_sound_speed(medium, c0)
    srcs = source isa AbstractVector ? source : [source]
    sens = sensor isa AbstractVector ? sensor : fill(sensor, length(srcs))
    obss = source isa AbstractVector ? p_obs : [p_obs]
    total = 0.0
    for (src, sen, obs) in zip(srcs, sens, obss)
        out = kspace_first_order(kgrid, med, src, sen; kwargs...)
        total += 0.5 * sum(abs2, out[:p] .- obs)
    end
    return total
end

end # module

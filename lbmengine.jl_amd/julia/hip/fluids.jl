#= hip backend of LBMengine.jl — fluids.jl

The fluid side of the reference's gpu/fluids.jl over liblbm_hip.so: the
device-held fields read through getproperty, hydroVariablesUpdate!,
LBMpropagate! (one model, or a vector of models advanced together),
viscousStressTensor and the population health check.  The time step itself
(tick!, gpu/fluids.jl:132-226) is lbm_run: collision, streaming, bounce-back,
moving walls, Guo forcing and the particles' momentum exchange and moves in
one kernel per tick. =#

using LinearAlgebra: norm

# ----------------------------------------------------------- device fields

gridSize(model::LBMmodel) = model.modelSize[1:model.spaceTime.dims]

function hydroFields(model::LBMmodel)
    N = gridSize(model)
    rho = zeros(N...)
    u = zeros(N..., model.spaceTime.dims)
    lbm_get_hydro!(getfield(model, :handle), rho, u)
    return rho, u
end

function populations(model::LBMmodel)
    f = zeros(gridSize(model)..., length(model.velocities.ws))
    lbm_get_populations!(getfield(model, :handle), f)
    return f
end

# massDensity, fluidVelocity, distributions, tick and time live on the
# device; every other field is the struct's own.
function Base.getproperty(model::LBMmodel, s::Symbol)
    s === :massDensity && return hydroFields(model)[1]
    s === :fluidVelocity && return hydroFields(model)[2]
    s === :distributions && return populations(model)
    s === :tick && return lbm_get_tick(getfield(model, :handle))[1]
    s === :time && return lbm_get_tick(getfield(model, :handle))[2]
    return getfield(model, s)
end

function Base.setproperty!(model::LBMmodel, s::Symbol, v)
    if s === :distributions
        f = Array{Float64}(v)
        @assert size(f) == (gridSize(model)..., length(model.velocities.ws)) "distributions of the wrong size"
        return lbm_set_populations(getfield(model, :handle), f)
    end
    @assert !(s in (:massDensity, :fluidVelocity, :tick, :time)) "$s is derived from the populations on the device"
    return setfield!(model, s, v)
end

Base.propertynames(::LBMmodel) = (:spaceTime, :tick, :time, :fluidParams, :initialConditions, :massDensity,
                                  :fluidVelocity, :forceDensity, :distributions, :velocities,
                                  :boundaryConditionsParams, :particles, :particleInteractions, :schemes)

# hydroVariablesUpdate! (gpu/fluids.jl:32-37): rho and u are computed from
# the populations whenever they are read, so there is nothing to refresh.
hydroVariablesUpdate!(model::LBMmodel; useEquilibriumScheme = false) = nothing

# --------------------------------------------------------------- the step

# the tick count of LBMpropagate! (gpu/fluids.jl:231-241)
function propagationTicks(Δt; simulationTime = :default, ticks = :default)
    @assert any(x -> x == :default, [simulationTime, ticks]) "simulationTime and ticks cannot be simultaneously chosen, as the time step is defined already in the model!"
    simulationTime == :default && ticks == :default && return 100
    ticks == :default && return length(range(Δt, stop = simulationTime, step = Δt))
    return ticks::Int
end

# After a run: a NaN or infinite population is an error (the run went
# unstable); populations at the 16-bit storage's clamp warn.
function guard(model::LBMmodel)
    bad, sat = lbm_check_populations(getfield(model, :handle))
    bad > 0 && error("$bad non-finite populations after tick $(model.tick): the run went unstable")
    sat > 0 && @warn "$sat populations at the f16c storage's clamp after tick $(model.tick): use storage = :f32 for this flow"
    return nothing
end

# `n` ticks on the device; a time-dependent magnetic field goes up, chunk
# by chunk, for the ticks each chunk runs.
function runTicks!(model::LBMmodel, n::Integer)
    h = getfield(model, :handle)
    fieldFn = scheduledField(model)
    if fieldFn === nothing
        lbm_run(h, n)
        return nothing
    end
    done = 0
    while done < n
        k = min(n - done, scheduleChunk)
        uploadSchedule!(model, fieldFn, k)
        lbm_run(h, k)
        done += k
    end
    return nothing
end

function LBMpropagate!(model::LBMmodel; simulationTime = :default, ticks = :default, verbose = false,
                       ticksSaved = :default)
    n = propagationTicks(model.spaceTime.timeStep; simulationTime = simulationTime, ticks = ticks)
    if :saveData in model.schemes
        propagateSaving!(model, n, ticksSaved)
    else
        runTicks!(model, n)
        lbm_sync(getfield(model, :handle))
    end
    guard(model)
    verbose && println("\r t = $(round(model.time; digits = 2))")
    return nothing
end

# Several independent models (one lattice, storage, collision and force
# kind; any sizes, parameters and particles) advanced together: every
# kernel of the run covers all of them.  Each model ends where
# LBMpropagate!(model; ticks) would leave it.  The reference loops
# LBMpropagate! over the models of a sweep instead.
function LBMpropagate!(models::Vector{LBMmodel}; simulationTime = :default, ticks = :default, verbose = false)
    @assert !any(m -> :saveData in m.schemes, models) "saveData runs one model at a time"
    Δts = unique([m.spaceTime.timeStep for m in models])
    @assert simulationTime == :default || length(Δts) == 1 "simulationTime needs one time step across the models; give ticks instead"
    n = propagationTicks(Δts[1]; simulationTime = simulationTime, ticks = ticks)
    for m in models
        f = scheduledField(m)
        f === nothing || uploadSchedule!(m, f, n)
    end
    b = lbm_batch_create(Ptr{Cvoid}[getfield(m, :handle) for m in models])
    try
        lbm_batch_autotune(b)
        lbm_batch_run(b, n)
        lbm_batch_sync(b)
    finally
        lbm_batch_destroy(b)
    end
    foreach(guard, models)
    verbose && println("\r t = $(round(models[1].time; digits = 2))")
    return nothing
end

"""
    saveState(model, path)
    loadState!(model, path)

Write the state a run resumes from (populations in the storage's encoding,
every particle's state record, the tick and the time) to `path`, and resume
a model built the same way from it, bit for bit.  No counterpart in the
reference, which cannot resume a run.
"""
saveState(model::LBMmodel, path::AbstractString) = write(path, lbm_save_state(getfield(model, :handle)))
function loadState!(model::LBMmodel, path::AbstractString)
    lbm_load_state(getfield(model, :handle), read(path))
    return model
end

# viscousStressTensor (cpu/fluids.jl:507-520) evaluated on the device: a
# dims x dims matrix of fields, as the reference returns it.
function viscousStressTensor(model::LBMmodel)
    d = model.spaceTime.dims
    sigma = zeros(gridSize(model)..., d, d)
    lbm_viscous_stress!(getfield(model, :handle), sigma)
    return [sigma[ntuple(_ -> Colon(), d)..., a, b] for a in 1:d, b in 1:d]
end

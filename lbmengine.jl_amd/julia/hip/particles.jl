#= hip backend of LBMengine.jl — particles.jl

The particle side of the reference's gpu/particles.jl and the output of
gpu/aux.jl:271-401 over liblbm_hip.so.  moveParticles! and
molecularDynamicsTick! run inside the step kernels; here are the particles'
read-through fields, the hand-over of model.particleInteractions, the
per-tick field of a time-dependent magnetic field, saveData, and the
gradient of a swimming-speed loss. =#

# ------------------------------------------------------------ particle view

function particleState(p::LBMparticle)
    s = lbm_get_particle(getfield(p.model, :handle), p.id)
    return s, p.model.spaceTime.dims
end

# position, velocity, angularVelocity, nodeVelocity's parameters and the
# loads live on the device
function Base.getproperty(p::LBMparticle, s::Symbol)
    if s in (:position, :velocity, :angularVelocity, :swimmingDirection)
        st, d = particleState(p)
        s === :position && return collect(st.position[1:d])
        s === :velocity && return collect(st.velocity[1:d])
        s === :swimmingDirection && return collect(st.swimming_direction[1:d])
        return d == 2 ? st.angular_velocity[1] : collect(st.angular_velocity)
    elseif s === :forceInput || s === :torqueInput
        F, T = lbm_get_particle_load(getfield(p.model, :handle), getfield(p, :id))
        d = p.model.spaceTime.dims
        return s === :forceInput ? F[1:d] : (d == 2 ? T[1] : T)
    elseif s === :boundaryConditionsParams
        node, dir = lbm_get_links(getfield(p.model, :handle), getfield(p, :id))
        return (; links = collect(zip(node .+ 1, dir .+ 1)))   # (node, direction), 1-based
    end
    return getfield(p, s)
end

# ---------------------------------------------------- interaction hand-over

# molecularDynamicsTick!'s terms (gpu/particles.jl:100-108) in its order:
# one per bond, one per pair of a pair list; a constant zero dipole field
# adds nothing (`norm(B) ≈ 0`).  Ids are 0-based on the C side.
function libraryTerms(it::AbstractInteraction, dims)
    ids3(a, b, c = 0) = (Int32(a - 1), Int32(b - 1), Int32(c == 0 ? 0 : c - 1))
    pad4(v...) = ntuple(k -> k <= length(v) ? Float64(v[k]) : 0.0, 4)
    kind = LBM_INTERACTION[typeof(it)]
    if it isa LinearInteraction
        return [LbmInteraction(kind, ids3(it.id1, it.id2), pad4(it.equilibriumDisplacement, it.stiffness), (0.0, 0.0, 0.0))]
    elseif it isa BistableInteraction
        return [LbmInteraction(kind, ids3(it.id1, it.id2), pad4(it.trapRadius, it.a, it.b), (0.0, 0.0, 0.0))]
    elseif it isa PolarInteraction
        return [LbmInteraction(kind, ids3(it.id1, it.id2, it.id3), pad4(it.equilibriumAngle, it.stiffness), (0.0, 0.0, 0.0))]
    elseif it isa DipoleDipoleInteraction
        B = it.magneticField
        if B isa Function
            # read tick by tick from the handle's schedule
            scalar = fieldValue(B, 0.0, dims)[2]
            return [LbmInteraction(kind, ids3(a, b), pad4(it.dipoleConstant, scalar ? 1 : 0, 1), (0.0, 0.0, 0.0))
                    for (a, b) in it.pairs]
        end
        Bv = B isa Number ? (Float64(B), 0.0, 0.0) : ntuple(k -> k <= length(B) ? Float64(B[k]) : 0.0, 3)
        all(iszero, Bv) && return LbmInteraction[]
        return [LbmInteraction(kind, ids3(a, b), pad4(it.dipoleConstant, B isa Number ? 1 : 0), Bv) for (a, b) in it.pairs]
    else
        return [LbmInteraction(kind, ids3(a, b), pad4(it.epsilon, it.sigma, it.cutoff), (0.0, 0.0, 0.0)) for (a, b) in it.pairs]
    end
end

function syncInteractions!(model::LBMmodel)
    terms = reduce(vcat, [libraryTerms(it, model.spaceTime.dims) for it in model.particleInteractions];
                   init = LbmInteraction[])
    lbm_set_interactions(getfield(model, :handle), terms)
    return nothing
end

# ------------------------------------------------ time-dependent magnetic field

const scheduleChunk = 1 << 16   # ticks per uploaded chunk (24 bytes a tick)

# (B, scalar) at time t: `magneticField(t) |> cu` (gpu/initMethods.jl:339)
# keeps a number and rounds a Float64 vector to Float32
function fieldValue(fn, t, dims)
    v = fn(t)
    v isa Number && return (Float64(v), 0.0, 0.0), true
    @assert length(v) == dims "magneticField must have $dims components"
    w = eltype(v) <: AbstractFloat ? Float64.(Float32.(v)) : Float64.(v)
    return ntuple(k -> k <= dims ? w[k] : 0.0, 3), false
end

function scheduledField(model::LBMmodel)
    for it in model.particleInteractions
        it isa DipoleDipoleInteraction && it.magneticField isa Function && return it.magneticField
    end
    return nothing
end

# the field of the next n ticks: tick k ends at the time tick! reaches by
# adding Δt k times (gpu/fluids.jl:212), the time molecularDynamicsTick! sees
function uploadSchedule!(model::LBMmodel, fn, n)
    h = getfield(model, :handle)
    tick, t = lbm_get_tick(h)
    B = zeros(3, n)
    for k in 1:n
        t += model.spaceTime.timeStep
        B[:, k] .= fieldValue(fn, t, model.spaceTime.dims)[1]
    end
    lbmcheck(ccall((:lbm_set_field_schedule, libhip), Cint, (Ptr{Cvoid}, Ptr{Float64}, Int64, Int64),
                   h, B, tick + 1, n))
end

# ------------------------------------------------------------------ saveData

const outputDir = "output.lbm"

# writeTrajectories (gpu/aux.jl:271-317) through the library's writer
function writeTrajectories(model::LBMmodel)
    d = model.spaceTime.dims
    rho, u = hydroFields(model)
    lbm_write_fluid_csv(joinpath(outputDir, "fluidTrj_$(model.tick).csv"), d, model.modelSize, model.tick,
                        model.time, model.spaceTime.coordinates, rho, u, length(model.velocities.ws),
                        populations(model))
end

# writeTensor (gpu/aux.jl:364-400): T a dims x dims matrix of fields
function writeTensor(model::LBMmodel, T, name::String)
    d = model.spaceTime.dims
    t = zeros(gridSize(model)..., d, d)
    for a in 1:d, b in 1:d
        t[ntuple(_ -> Colon(), d)..., a, b] .= T[a, b]
    end
    mkpath(outputDir)
    lbm_write_tensor_csv(joinpath(outputDir, "$name.csv"), d, model.modelSize, model.tick, model.time,
                         model.spaceTime.coordinates, t)
end

function startSaving!(model::LBMmodel)
    rm(outputDir; force = true, recursive = true)
    mkdir(outputDir)
    writeTrajectories(model)
end

particleHeader(d) = join(vcat(["tick", "time", "particleId"], ["position_$a" for a in "xyz"[1:d]],
                              ["velocity_$a" for a in "xyz"[1:d]],
                              d == 2 ? ["angularVelocity"] : ["angularVelocity_$a" for a in "xyz"]), ",")

function appendParticleRows(model::LBMmodel, rows)
    path = joinpath(outputDir, "particlesTrj.csv")
    new = !isfile(path)
    open(path, "a") do io
        new && println(io, particleHeader(model.spaceTime.dims))
        foreach(r -> println(io, r), rows)
    end
end

particleRow(d, tick, time, id, R, V, W) =
    join(string.(vcat([tick, time, id], R[1:d], V[1:d], d == 2 ? [W[1]] : W[1:3])), ",")

# writeParticleTrajectory (gpu/aux.jl:319-356) of particle id
function writeParticle(model::LBMmodel, id)
    p = model.particles[id]
    appendParticleRows(model, [particleRow(model.spaceTime.dims, model.tick, model.time, id, p.position,
                                           p.velocity, collect(p.angularVelocity))])
end

# LBMpropagate! with :saveData (gpu/fluids.jl:243-259): fluid files at the
# checkpoint ticks, a particle row per particle per tick, read from the
# device's per-tick log after each stretch of at most logTicks ticks
const logTicks = 4096
function propagateSaving!(model::LBMmodel, ticks, ticksSaved)
    ticksSaved == :default && (ticksSaved = 100)
    h = getfield(model, :handle)
    t0 = model.tick
    Δt = model.spaceTime.timeStep
    total = floor(range(Δt, length = ticks, step = Δt)[end] / Δt)
    ckpt = Set(Int64.(round.(collect(range(t0, stop = t0 + total, length = ticksSaved)))))
    np = length(model.particles)
    d = model.spaceTime.dims
    np > 0 && lbm_set_trajectory_log(h, min(max(ticks, 1), logTicks))
    try
        done = 0
        while done < ticks
            a = t0 + done
            b = np > 0 ? min(t0 + ticks, a + logTicks) : t0 + ticks
            nxt = [c for c in ckpt if a < c <= b]
            isempty(nxt) || (b = minimum(nxt))
            ta = model.time
            runTicks!(model, b - a)
            lbm_sync(h)
            if np > 0
                logs = [lbm_get_trajectory(h, p, a + 1, b - a) for p in 1:np]
                rows, t = String[], ta
                for i in 1:(b - a)
                    t += Δt
                    for p in 1:np
                        e = logs[p][:, i]
                        push!(rows, particleRow(d, a + i, t, p, e[1:3], e[4:6], e[7:9]))
                    end
                end
                appendParticleRows(model, rows)
            end
            b in ckpt && writeTrajectories(model)
            done = b - t0
        end
    finally
        np > 0 && lbm_set_trajectory_log(h, 0)
    end
end

# --------------------------------------------------------------- gradients

# The loss L = (mean over the last `window` ticks of the swimmer's speed
# |V| - targetSpeed)^2 + forcePenalty * sum(F^2) over `ticks` ticks from the
# model's current populations, with its gradient along every particle's B1,
# B2 and (with a per-node force field) the force-density field, by the discrete adjoint of the step on
# the device (checkpoints every `checkpointEvery` ticks).  The model ends
# `ticks` ahead.  No counterpart in the reference, whose step is not
# differentiable.
function swimmingLossGradient(model::LBMmodel; ticks, checkpointEvery = 100, window = ticks, targetSpeed = 1e-4,
                              forcePenalty = 0.0, particle = 1)
    cfg = LbmAdjointConfig(ticks, checkpointEvery, window, particle - 1, 0, targetSpeed, forcePenalty)
    # the field gradient needs a per-node force field
    dF = model.forceKind == LBM_FORCE_FIELD ? zeros(gridSize(model)..., model.spaceTime.dims) : nothing
    L, dB1, dB2 = lbm_adjoint_run(getfield(model, :handle), cfg, length(model.particles), dF)
    return L, (; B1 = dB1, B2 = dB2, forceDensity = dF)
end

# The same loss by a plain forward run: the swimmer's velocities of the last
# `window` ticks come from the device's per-tick log.
function swimmingLoss(model::LBMmodel; ticks, window = ticks, targetSpeed = 1e-4, forcePenalty = 0.0,
                      particle = 1)
    h = getfield(model, :handle)
    runTicks!(model, ticks - window)
    lbm_set_trajectory_log(h, window)
    try
        first = model.tick + 1
        runTicks!(model, window)
        lbm_sync(h)
        log = lbm_get_trajectory(h, particle, first, window)
        d = model.spaceTime.dims
        speed = sum(norm(log[4:(3 + d), k]) for k in 1:window) / window
        L = (speed - targetSpeed)^2
        model.forceKind == LBM_FORCE_FIELD && (L += forcePenalty * sum(abs2, model.forceDensity))
        return L
    finally
        lbm_set_trajectory_log(h, 0)
    end
end

#= hip backend of LBMengine.jl — initMethods.jl

modelInit, addBead!, addSquirmer! and the particle-interaction constructors
with the reference's keywords, defaults and assertions
(gpu/initMethods.jl:7-414, 440-702).  The unit, scheme and initial-condition
resolution is host code and stays here; the device allocation, the initial
populations and the particles' first placement are liblbm_hip.so calls. =#

# ----------------------------------------------------------------- modelInit

function modelInit(;
    massDensity = :default, fluidVelocity = :default, velocities = :default, collisionModel = :default,
    relaxationTimeRatio = :default, viscosity = :default, timeStep = :default, latticeParameter = :default,
    kineticParameter = 0.25, xlims = :default, ylims = :default, zlims = :default, dims = 2,
    walledDimensions = :default, dampenEcho = false, walls = :default, movingWalls = :default,
    isFluidCompressible = true, forceDensity = :default, forcingScheme = :default, saveData = false,
    # backend keywords: population storage, which reference backend to follow, deviation shift
    storage = :f64, semantics = :gpu, rho0 = 1.0,
)
    # time step, lattice parameter, relaxation time and viscosity (:469-492)
    @assert any(x -> x == :default, [timeStep, relaxationTimeRatio, viscosity]) "timeStep, relaxationTimeRatio, and viscosity cannot be all simultaneously defined!"
    @assert any(x -> x == :default, [timeStep, latticeParameter]) "timeStep, and latticeParameter cannot be both simultaneously defined!"
    if latticeParameter != :default
        @assert latticeParameter isa Number "latticeParameter must be a number!"
        timeStep = latticeParameter
    end
    if timeStep == :default
        relaxationTimeRatio == :default && (relaxationTimeRatio = 15)
        viscosity == :default && (viscosity = 1)
        timeStep = 3 * viscosity / (relaxationTimeRatio - 0.5)
    elseif relaxationTimeRatio == :default
        viscosity == :default && (viscosity = 1)
        relaxationTimeRatio = 3 * viscosity / timeStep + 0.5
    else
        viscosity = timeStep / 3 * (relaxationTimeRatio - 0.5)
    end
    Δx = Δt = timeStep

    # coordinates (:495-505)
    xlims == :default && (xlims = (-1, 1))
    ylims == :default && (ylims = xlims)
    zlims == :default && (zlims = xlims)
    @assert all(t -> t[2] != t[1], [xlims, ylims, zlims]) "Coordinate lower and upper limits cannot be equal!"
    coordinates = [coordinateRange(l, Δx) for l in (xlims, ylims, zlims)][1:dims]
    N = Tuple(length.(coordinates))

    @assert !dampenEcho ":cbc is not implemented in the gpu modules"
    forcingScheme == :default && (forcingScheme = :guo)
    @assert forcingScheme == :guo "forcing scheme $forcingScheme is not implemented in the hip backend"

    # initial hydrodynamic fields (:507-527)
    if massDensity == :default
        ρ = 1.0
    elseif massDensity isa Number
        ρ = Float64(massDensity)
    else
        @assert size(massDensity) == N "The mass density given has the wrong size! $N expected."
        ρ = Array{Float64}(massDensity)
    end
    if fluidVelocity == :default
        u = zeros(dims)
    elseif ndims(fluidVelocity) == 1
        u = Vector{Float64}(fluidVelocity)
    else
        @assert size(fluidVelocity)[1:dims] == N "The fluid velocity given has the wrong size! $N expected."
        u = Array{Float64}(fluidVelocity)
    end

    # velocity set (:529-541)
    if velocities == :default
        @assert dims > 1 "1 dimensional LBM is not implemented in the hip backend!"
        velocities = (:D1Q3, :D2Q9, :D3Q15)[dims]
    end
    vs = velocitySet(velocities)
    @assert length(vs.cs[1]) == dims "The chosen velocity must be $dims-dimensional!"

    # fluid parameters (:543-570)
    latticeSpeed = Float64(Δx / Δt)
    c_s, c2_s, c4_s = latticeSpeed / √3, latticeSpeed^2 / 3, latticeSpeed^4 / 9
    invC_s, invC2_s, invC4_s = √3 / latticeSpeed, 3 / latticeSpeed^2, 9 / latticeSpeed^4
    collisionModel == :default && (collisionModel = :trt)
    @assert collisionModel in (:bgk, :trt) "Collision model $collisionModel is not implemented!"
    relaxationTime = relaxationTimeRatio * Δt
    schemes = Symbol[collisionModel]
    τminus = NaN
    if collisionModel == :bgk
        fluidParams = (; c_s, c2_s, c4_s, invC_s, invC2_s, invC4_s, relaxationTime, viscosity, isFluidCompressible)
    else
        Λ = kineticParameter == :debug ? viscosity^2 / (c4_s * Δt^2) : kineticParameter
        τminus = Λ * c2_s * Δt^2 / viscosity + Δt / 2
        fluidParams = (; c_s, c2_s, c4_s, invC_s, invC2_s, invC4_s, relaxationTimePlus = relaxationTime,
                       relaxationTimeMinus = τminus, viscosity, isFluidCompressible)
    end

    # walls: the walled layers, the `walls` region, the moving walls (:572-626)
    walled = Int32(0)
    if walledDimensions != :default && !isempty(walledDimensions)
        for k in walledDimensions
            @assert 1 <= k <= dims "walledDimensions are 1-based dimensions of the model"
            walled |= Int32(1) << (k - 1)
        end
        push!(schemes, :bounceBack)
    end
    wallRegion = wallNodes(N, walledDimensions == :default ? () : walledDimensions)
    wallMask = nothing
    if walls != :default
        @assert walls isa Function "walls must be a function!"
        wallMask = [walls(x...) for x in gridPoints(coordinates)]
        @assert eltype(wallMask) == Bool "walls must return boolean values!"
        wallRegion .|= wallMask
        push!(schemes, :bounceBack)
    end
    wallVelocity = nothing
    if movingWalls != :default
        @assert movingWalls isa Function "movingWalls must be a function!"
        vals = [Vector{Float64}(movingWalls(x...)) for x in gridPoints(coordinates)]
        @assert length(vals[1]) == dims "movingWalls must return a vector of dimension $dims!"
        wallVelocity = zeros(N..., dims)
        for (i, v) in enumerate(vals), k in 1:dims
            wallVelocity[i + (k - 1) * prod(N)] = v[k]
        end
        wallRegion .|= [norm(v) > 0 for v in vals]
        append!(schemes, [:movingWalls, :bounceBack])
    end

    # the incompressible equilibrium w ρ + w ρ0 (...) (gpu/fluids.jl:51), ρ0 uniform
    flags = Int32(0)
    if !isFluidCompressible
        @assert ρ isa Number "the incompressible equilibrium needs a uniform initial density in the hip backend"
        flags |= LBM_FLAG_INCOMPRESSIBLE
        rho0 = ρ
    end

    # body force (:628-647)
    force, Funiform, Ffield = LBM_FORCE_NONE, (0.0, 0.0, 0.0), nothing
    if forceDensity != :default
        if ndims(forceDensity) == 1 && length(forceDensity) == dims
            force = LBM_FORCE_UNIFORM
            Funiform = ntuple(k -> k <= dims ? Float64(forceDensity[k]) : 0.0, 3)
        elseif size(forceDensity) == (N..., dims)
            force = LBM_FORCE_FIELD
            Ffield = Array{Float64}(forceDensity)
        else
            error("force density does not have consistent dimensions!")
        end
        push!(schemes, forcingScheme)
    end
    saveData && push!(schemes, :saveData)

    @assert haskey(LBM_STORAGE, storage) "storage must be one of $(collect(keys(LBM_STORAGE)))"
    @assert haskey(LBM_SEMANTICS, semantics) "semantics must be one of $(collect(keys(LBM_SEMANTICS)))"
    n3 = ntuple(k -> k <= dims ? Int64(N[k]) : Int64(1), 3)
    cfg = LbmConfig(dims, length(vs.ws), n3, LBM_COLLISION[collisionModel], LBM_STORAGE[storage],
                    LBM_SEMANTICS[semantics], walled, force, flags, Funiform, Δt, Δx, relaxationTime,
                    τminus, Float64(rho0),
                    ntuple(k -> k <= dims ? pointer(coordinates[k]) : Ptr{Float64}(C_NULL), 3))
    h = GC.@preserve coordinates lbm_create(cfg)

    Fstored = force == LBM_FORCE_FIELD ? Ffield :
              force == LBM_FORCE_UNIFORM ? repeat(reshape(collect(Funiform[1:dims]), ntuple(_ -> 1, dims)..., dims), N..., 1) :
              zeros(N..., dims)
    spaceTime = (; coordinates, latticeParameter = Δx, timeStep = Δt, latticeSpeed, dims)
    model = LBMmodel(h, spaceTime, fluidParams, (; massDensity = ρ), (; cs = vs.cs, ws = vs.ws),
                     (; wallRegion, oppositeVectorId = oppositeVectorId(vs.cs)), LBMparticle[], AbstractInteraction[],
                     schemes, Fstored, force, (N..., ntuple(_ -> 1, 3 - dims)...), Any[coordinates])
    finalizer(m -> lbm_destroy(getfield(m, :handle)), model)

    # walls before the populations; the library adds the walled layers and the
    # moving walls' nodes to the region itself
    wallMask === nothing || lbm_set_walls(h, Array{UInt8}(wallMask))
    wallVelocity === nothing || lbm_set_wall_velocity(h, wallVelocity)
    force == LBM_FORCE_FIELD && lbm_set_force_field(h, Ffield)

    # findInitialConditions (:417-438, 657-668)
    if ρ isa Number && ndims(u) == 1
        u3 = zeros(3); u3[1:dims] = u
        lbm_init_uniform(h, ρ, u3)
    else
        lbm_set_populations(h, initialPopulations(model, ρ, u, wallRegion))
    end
    saveData && startSaving!(model)
    return model
end

# findInitialConditions (gpu/initMethods.jl:417-438) for non-uniform fields:
# the equilibrium at u - Δt F / (2ρ), zero on walls; (N..., Q)
function initialPopulations(model::LBMmodel, ρ, u, wallRegion)
    d = model.spaceTime.dims
    N = gridSize(model)
    ρf = ρ isa Number ? fill(ρ, N...) : copy(ρ)
    uf = ndims(u) == 1 ? repeat(reshape(u, ntuple(_ -> 1, d)..., d), N..., 1) : copy(u)
    ρf[wallRegion] .= 0
    for k in 1:d
        selectdim(uf, d + 1, k)[wallRegion] .= 0
    end
    F = copy(model.forceDensity)
    for k in 1:d
        selectdim(F, d + 1, k)[wallRegion] .= 0
        Fk, uk = selectdim(F, d + 1, k), selectdim(uf, d + 1, k)
        uk .-= (0.5 * model.spaceTime.timeStep) .* ifelse.(ρf .!= 0, Fk ./ ρf, 0.0)
    end
    p = model.fluidParams
    f = zeros(N..., length(model.velocities.ws))
    for (q, (c, w)) in enumerate(zip(model.velocities.cs, model.velocities.ws))
        ci = c .* model.spaceTime.latticeSpeed
        udc = sum(selectdim(uf, d + 1, k) .* ci[k] for k in 1:d)
        uu = sum(selectdim(uf, d + 1, k) .^ 2 for k in 1:d)
        first = udc .* p.invC2_s .+ (udc .* udc) .* (0.5 * p.invC4_s)
        selectdim(f, d + 1, q) .= (first .- uu .* (0.5 * p.invC2_s) .+ 1.0) .* (w .* ρf)
    end
    return f
end

# ------------------------------------------------------------------ particles

# sum(getSphere(radius, X)): nodes strictly inside the origin-centred sphere
sphereCount(model::LBMmodel, radius) =
    count(x -> sum(abs2, x) < radius^2, gridPoints(model.spaceTime.coordinates))

function particleRecord(model::LBMmodel, kind; massDensity = 1.0, radius = 0.1, position = :default,
                        velocity = :default, angularVelocity = :default, coupleTorques = false,
                        coupleForces = false, scheme = :default)
    d = model.spaceTime.dims
    mass = massDensity * sphereCount(model, radius) * model.spaceTime.latticeParameter^d
    position == :default && (position = zeros(d))
    velocity == :default && (velocity = zeros(d))
    @assert length(position) == length(velocity) == d "The position and velocity dimensions must match the dimensionality of the model! dims = $d"
    if d == 2
        I = 0.5 * mass * radius^2
        angularVelocity == :default && (angularVelocity = 0.0)
        @assert angularVelocity isa Number "In two dimensions, angularVelocity must be a number!"
        ω = (Float64(angularVelocity), 0.0, 0.0)
    else
        I = 0.4 * mass * radius^2
        angularVelocity == :default && (angularVelocity = zeros(3))
        @assert length(angularVelocity) == 3 "In three dimensions, angularVelocity must be an array!"
        ω = Tuple(Float64.(angularVelocity))
    end
    scheme == :default && (scheme = :ladd)
    @assert scheme in (:psm, :ladd) "$scheme cannot be used as a particle-fluid collision scheme!"
    @assert scheme == :ladd "$scheme is not implemented in the gpu modules!"
    pad(v) = ntuple(k -> k <= d ? Float64(v[k]) : 0.0, 3)
    return LbmParticle(kind, coupleForces, coupleTorques, 0, radius, pad(position), pad(velocity), ω,
                       (0.0, 0.0, 0.0), 0.0, 0.0, 1 / mass, 1 / I)
end

function registerParticle!(model::LBMmodel, p::LbmParticle, params)
    id = lbm_add_particle(getfield(model, :handle), p)
    push!(model.particles, LBMparticle(model, id, params))
    for s in (:ladd, :bounceBack)
        s in model.schemes || push!(model.schemes, s)
    end
    :saveData in model.schemes && writeParticle(model, id)
    return nothing
end

# addBead! (gpu/initMethods.jl:7-89)
function addBead!(model::LBMmodel; kw...)
    p = particleRecord(model, LBM_PARTICLE_BEAD; kw...)
    registerParticle!(model, p, (; radius = p.radius, properties = [:spherical, :bead],
                                 coupleForces = p.couple_forces != 0, coupleTorques = p.couple_torques != 0))
end

# addSquirmer! (gpu/initMethods.jl:91-222)
function addSquirmer!(model::LBMmodel; slipSpeed = :default, swimmingSpeed = :default,
                      swimmingDirection = :default, B1 = :default, B2 = :default, beta = :default, kw...)
    d = model.spaceTime.dims
    p = particleRecord(model, LBM_PARTICLE_SQUIRMER; kw...)
    @assert any(x -> x == :default, [B1, B2, beta]) "B1, B2, and beta cannot be all simultaneously defined!"
    if beta == :default
        B1 == :default && (B1 = 1)
        B2 == :default && (B2 = 0)
        beta = B2 / B1
    elseif B2 == :default
        B1 == :default && (B1 = 1)
        B2 = B1 * beta
    else
        @assert beta != 0 "if beta = 0, then B2 can only be zero! In this case, do not declare both."
        B1 = B2 / beta
    end
    swimmingDirection == :default && (swimmingDirection = [1.0; zeros(d - 1)])
    @assert length(swimmingDirection) == d "the swimming direction must be $d-dimensional!"
    e = Vector{Float64}(swimmingDirection)
    e = e / norm(e)
    @assert any(x -> x == :default, [slipSpeed, swimmingSpeed]) "slipSpeed and swimmingSpeed cannot be all simultaneously defined!"
    if slipSpeed == :default
        swimmingSpeed == :default && (swimmingSpeed = 1e-3)
        e = e * sign(swimmingSpeed)
        B1 = abs(3 / 2 * swimmingSpeed)
        B2 = B1 * beta
    else
        θs = range(-π, stop = π, length = 100)
        peak = maximum(θ -> abs((B1 + B2 * cos(θ)) * sin(θ)), θs)
        B1, B2 = B1 * slipSpeed / peak, B2 * slipSpeed / peak
    end
    q = LbmParticle(p.kind, p.couple_forces, p.couple_torques, 0, p.radius, p.position, p.velocity,
                    p.angular_velocity, ntuple(k -> k <= d ? e[k] : 0.0, 3), Float64(B1), Float64(B2),
                    p.inverse_mass, p.inverse_moment_of_inertia)
    registerParticle!(model, q, (; radius = q.radius, properties = [:spherical, :squirmer], B1 = q.B1, B2 = q.B2,
                                 coupleForces = q.couple_forces != 0, coupleTorques = q.couple_torques != 0))
end

# --------------------------------------------------- particle interactions

checkIds(model, ids...) = foreach(i -> @assert(1 <= i <= length(model.particles), "there is no particle $i"), ids)

# addLinearBond! (gpu/initMethods.jl:229-245)
function addLinearBond!(model::LBMmodel, id1::Int64, id2::Int64; equilibriumDisplacement = :default, stiffness = 1)
    checkIds(model, id1, id2)
    equilibriumDisplacement == :default &&
        (equilibriumDisplacement = norm(model.particles[id2].position - model.particles[id1].position))
    @assert equilibriumDisplacement isa Number "equilibriumDisplacement must be a number!"
    @assert stiffness isa Number "stiffness must be a number!"
    push!(model.particleInteractions, LinearInteraction(id1, id2, equilibriumDisplacement, stiffness))
    syncInteractions!(model)
end

# addPolarBond! (gpu/initMethods.jl:247-273)
function addPolarBond!(model::LBMmodel, id1::Int64, id2::Int64, id3::Int64; equilibriumAngle = :default, stiffness = 1)
    checkIds(model, id1, id2, id3)
    a = model.particles[id1].position - model.particles[id2].position
    b = model.particles[id3].position - model.particles[id2].position
    equilibriumAngle == :default && (equilibriumAngle = acos(sum(a .* b) / (norm(a) * norm(b))))
    @assert equilibriumAngle isa Number "equilibriumAngle must be a number!"
    @assert stiffness isa Number "stiffness must be a number!"
    push!(model.particleInteractions, PolarInteraction(id1, id2, id3, equilibriumAngle, stiffness))
    syncInteractions!(model)
end

# addBistableBond! (gpu/initMethods.jl:282-308): a quartic double well with
# minima at lowDisp and highDisp and a barrier hillHeight between them
function addBistableBond!(model::LBMmodel, id1::Int64, id2::Int64; lowDisp = :default, highDisp = :default, hillHeight = 1)
    checkIds(model, id1, id2)
    lowDisp == :default && (lowDisp = norm(model.particles[id2].position - model.particles[id1].position))
    highDisp == :default && (highDisp = lowDisp + model.spaceTime.latticeParameter)
    trapRadius, trapWidth = (highDisp + lowDisp) / 2, Float64(highDisp - lowDisp)
    a, b = 4 * 16 * hillHeight / trapWidth^4, 2 * 8 * hillHeight / trapWidth^2
    push!(model.particleInteractions, BistableInteraction(id1, id2, trapRadius, trapWidth, hillHeight, a, b))
    syncInteractions!(model)
end

# addDipoles! (gpu/initMethods.jl:316-359): every pair of the given particles
# (all when none is given); a later call adds its ids unless rewrite.  The
# field is a vector, in 2D a number, or a function of time returning either.
function addDipoles!(model::LBMmodel, ids...; magneticField = [1, 0], susceptibility = 0.5,
                     permeability = 400 * pi, rewrite = false)
    ids = isempty(ids) ? collect(eachindex(model.particles)) : collect(ids)
    checkIds(model, ids...)
    at = findfirst(i -> i isa DipoleDipoleInteraction, model.particleInteractions)
    writeNew = at === nothing || rewrite
    if !writeNew
        ids = sort(unique(vcat(ids, [collect(p) for p in model.particleInteractions[at].pairs]...)))
    end
    @assert length(ids) > 1 "There must be at least two dipoles!"
    pairs = Tuple{UInt8,UInt8}[(ids[i], ids[j]) for i in eachindex(ids) for j in eachindex(ids) if i < j]
    if writeNew
        rewrite && at !== nothing && deleteat!(model.particleInteractions, at)
        r = model.particles[ids[1]].particleParams.radius
        push!(model.particleInteractions,
              DipoleDipoleInteraction(pairs, 4 * pi * r^6 * susceptibility^2 / (3 * permeability), magneticField))
    else
        model.particleInteractions[at].pairs = pairs
    end
    syncInteractions!(model)
end

# addLennardJones! (gpu/initMethods.jl:361-396) and addWCA! (:398-414)
function defaultEpsilonSigma(model, epsilon, sigma)
    Δx, Δt = model.spaceTime.latticeParameter, model.spaceTime.timeStep
    if epsilon == :default
        ρ = filter(!iszero, vec(model.massDensity))
        epsilon = sum(ρ) / length(ρ) * Δx^5 / Δt^2
    end
    sigma == :default && (sigma = 2 * model.particles[1].particleParams.radius + Δx)
    return epsilon, sigma
end

function addLennardJones!(model::LBMmodel; epsilon = :default, sigma = :default, cutoff = :default)
    at = findfirst(i -> i isa LennardJonesInteraction, model.particleInteractions)
    at === nothing || deleteat!(model.particleInteractions, at)
    ids = collect(eachindex(model.particles))
    pairs = Tuple{UInt8,UInt8}[(ids[i], ids[j]) for i in eachindex(ids) for j in eachindex(ids) if i < j]
    epsilon, sigma = defaultEpsilonSigma(model, epsilon, sigma)
    cutoff == :default && (cutoff = Inf)
    push!(model.particleInteractions, LennardJonesInteraction(pairs, epsilon, sigma, cutoff))
    syncInteractions!(model)
end

function addWCA!(model::LBMmodel; epsilon = :default, sigma = :default)
    epsilon, sigma = defaultEpsilonSigma(model, epsilon, sigma)
    addLennardJones!(model; epsilon = epsilon, sigma = sigma, cutoff = 2^(1 / 6) * sigma)
end

#= hip backend of LBMengine.jl — structs.jl

The types this backend's functions take and return.  `LBMmodel` keeps the
reference's field names (gpu/structs.jl:10-33); the fields that live on the
device (massDensity, fluidVelocity, distributions, tick, time) are read
through `getproperty` from liblbm_hip.so (fluids.jl of this backend), so
code written against the reference's model reads the same names.

The C structs mirror include/lbm_hip.h field for field. =#

# include/lbm_hip.h: lbm_config
struct LbmConfig
    dims::Int32
    q::Int32
    n::NTuple{3,Int64}
    collision::Int32
    storage::Int32
    semantics::Int32
    walled::Int32
    force::Int32
    flags::Int32
    force_uniform::NTuple{3,Float64}
    dt::Float64
    dx::Float64
    tau::Float64
    tau_minus::Float64
    rho0::Float64
    coords::NTuple{3,Ptr{Float64}}
end

# include/lbm_hip.h: lbm_particle
struct LbmParticle
    kind::Int32
    couple_forces::Int32
    couple_torques::Int32
    reserved0::Int32
    radius::Float64
    position::NTuple{3,Float64}
    velocity::NTuple{3,Float64}
    angular_velocity::NTuple{3,Float64}
    swimming_direction::NTuple{3,Float64}
    B1::Float64
    B2::Float64
    inverse_mass::Float64
    inverse_moment_of_inertia::Float64
end

# include/lbm_hip.h: lbm_adjoint_config
struct LbmAdjointConfig
    ticks::Int64
    checkpoint_every::Int64
    window::Int64
    particle::Int32
    reserved::Int32
    target_speed::Float64
    force_penalty::Float64
end

const LBM_COLLISION = Dict(:bgk => Int32(0), :trt => Int32(1))
const LBM_STORAGE = Dict(:f64 => Int32(0), :f32 => Int32(1), :f16 => Int32(2), :f16c => Int32(3))
const LBM_SEMANTICS = Dict(:gpu => Int32(0), :cpu => Int32(1))
const LBM_FORCE_NONE, LBM_FORCE_UNIFORM, LBM_FORCE_FIELD = Int32(0), Int32(1), Int32(2)
const LBM_FLAG_INCOMPRESSIBLE = Int32(1)
const LBM_PARTICLE_BEAD, LBM_PARTICLE_SQUIRMER = Int32(0), Int32(1)

# A lattice on the GPU: the reference's LBMmodel fields, host-side ones as
# plain fields, device-side ones behind getproperty (see fluids.jl).
mutable struct LBMmodel
    handle::Ptr{Cvoid}
    spaceTime::NamedTuple          # Δx, Δt, coordinates, latticeSpeed, dims (gpu/initMethods.jl:501-507)
    fluidParams::NamedTuple        # c_s powers, relaxation times, viscosity, isFluidCompressible
    initialConditions::NamedTuple  # (; massDensity = ρ₀)
    velocities::NamedTuple         # (; cs, ws)
    boundaryConditionsParams::NamedTuple
    particles::Vector{Any}
    particleInteractions::Vector{Any}
    schemes::Vector{Symbol}
    forceDensity::AbstractArray{Float64}
    forceKind::Int32               # LBM_FORCE_NONE, _UNIFORM or _FIELD
    modelSize::NTuple{3,Int}
    keepalive::Vector{Any}         # coordinate vectors the config points to
end

# One particle of a model (gpu/structs.jl:43-53): its state lives on the
# device and is read through getproperty (particles.jl).
struct LBMparticle
    model::LBMmodel
    id::Int                        # 1-based, as the reference's
    particleParams::NamedTuple
end

# model.particleInteractions entries (gpu/structs.jl:55-93): the same
# types and field names, since user code builds and reads them.  Ids are
# 1-based.  A DipoleDipoleInteraction's magneticField is a vector, a number
# (2D, a field along z) or a function of time returning either.
abstract type AbstractInteraction end

struct LinearInteraction <: AbstractInteraction
    id1::UInt8; id2::UInt8
    equilibriumDisplacement::Float64; stiffness::Float64
end

struct BistableInteraction <: AbstractInteraction
    id1::UInt8; id2::UInt8
    trapRadius::Float64; trapWidth::Float64; hillHeight::Float64; a::Float64; b::Float64
end

struct PolarInteraction <: AbstractInteraction
    id1::UInt8; id2::UInt8; id3::UInt8
    equilibriumAngle::Float64; stiffness::Float64
end

mutable struct DipoleDipoleInteraction <: AbstractInteraction
    pairs::Vector{Tuple{UInt8,UInt8}}
    dipoleConstant::Float64
    magneticField
end

mutable struct LennardJonesInteraction <: AbstractInteraction
    pairs::Vector{Tuple{UInt8,UInt8}}
    epsilon::Float64; sigma::Float64; cutoff::Float64
end

# include/lbm_hip.h: lbm_interaction
struct LbmInteraction
    kind::Int32
    ids::NTuple{3,Int32}
    params::NTuple{4,Float64}
    field::NTuple{3,Float64}
end
const LBM_INTERACTION = Dict(LinearInteraction => Int32(0), BistableInteraction => Int32(1),
                             PolarInteraction => Int32(2), DipoleDipoleInteraction => Int32(3),
                             LennardJonesInteraction => Int32(4))

#= hip backend of LBMengine.jl — aux.jl

The ccall layer over liblbm_hip.so (include/lbm_hip.h).  Every entry point
returns 0 or a negative LBM_E* code; `lbmcheck` turns a failure into a Julia
error carrying the library's message.  Host arrays are column-major Julia
arrays in the reference's layout; q is 0-based on the C side. =#

# (libhip, the library's path, and hipFunctional() come from probe.jl, which
# the module includes before it picks the backend)

lbmcheck(rc) = rc == 0 || error("liblbm_hip: ", unsafe_string(ccall((:lbm_last_error, libhip), Cstring, ())))

function lbm_create(cfg::LbmConfig)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    lbmcheck(ccall((:lbm_create, libhip), Cint, (Ref{LbmConfig}, Ref{Ptr{Cvoid}}), cfg, h))
    return h[]
end
lbm_destroy(h) = ccall((:lbm_destroy, libhip), Cint, (Ptr{Cvoid},), h)

lbm_init_uniform(h, rho::Float64, u::Vector{Float64}) =
    lbmcheck(ccall((:lbm_init_uniform, libhip), Cint, (Ptr{Cvoid}, Float64, Ptr{Float64}), h, rho, u))
lbm_set_force_field(h, F::Array{Float64}) =                  # (N..., dims)
    lbmcheck(ccall((:lbm_set_force_field, libhip), Cint, (Ptr{Cvoid}, Ptr{Float64}), h, F))
lbm_set_walls(h, wall::Array{UInt8}) =                        # (N...)
    lbmcheck(ccall((:lbm_set_walls, libhip), Cint, (Ptr{Cvoid}, Ptr{UInt8}), h, wall))
lbm_set_wall_velocity(h, u::Array{Float64}) =                 # (N..., dims)
    lbmcheck(ccall((:lbm_set_wall_velocity, libhip), Cint, (Ptr{Cvoid}, Ptr{Float64}), h, u))
lbm_set_populations(h, f::Array{Float64}) =                   # (N..., Q)
    lbmcheck(ccall((:lbm_set_populations, libhip), Cint, (Ptr{Cvoid}, Ptr{Float64}), h, f))
lbm_get_populations!(h, f::Array{Float64}) =
    lbmcheck(ccall((:lbm_get_populations, libhip), Cint, (Ptr{Cvoid}, Ptr{Float64}), h, f))
lbm_get_hydro!(h, rho, u) =
    lbmcheck(ccall((:lbm_get_hydro, libhip), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}), h,
                   rho === nothing ? C_NULL : rho, u === nothing ? C_NULL : u))

function lbm_add_particle(h, p::LbmParticle)
    id = Ref{Int32}(0)
    lbmcheck(ccall((:lbm_add_particle, libhip), Cint, (Ptr{Cvoid}, Ref{LbmParticle}, Ref{Int32}), h, p, id))
    return Int(id[]) + 1                                       # 1-based
end
function lbm_get_particle(h, id::Int)
    p = Ref{LbmParticle}()
    lbmcheck(ccall((:lbm_get_particle, libhip), Cint, (Ptr{Cvoid}, Int32, Ref{LbmParticle}), h, id - 1, p))
    return p[]
end
function lbm_get_particle_load(h, id::Int)
    F, T = zeros(3), zeros(3)
    lbmcheck(ccall((:lbm_get_particle_load, libhip), Cint, (Ptr{Cvoid}, Int32, Ptr{Float64}, Ptr{Float64}),
                   h, id - 1, F, T))
    return F, T
end

lbm_run(h, n::Integer) = lbmcheck(ccall((:lbm_run, libhip), Cint, (Ptr{Cvoid}, Int64), h, n))
lbm_sync(h) = lbmcheck(ccall((:lbm_sync, libhip), Cint, (Ptr{Cvoid},), h))
function lbm_get_tick(h)
    t, tm = Ref{Int64}(0), Ref{Float64}(0.0)
    lbmcheck(ccall((:lbm_get_tick, libhip), Cint, (Ptr{Cvoid}, Ref{Int64}, Ref{Float64}), h, t, tm))
    return t[], tm[]
end
# state snapshots (no counterpart in the reference): the bytes a run resumes from
function lbm_save_state(h)
    n = Ref{Int64}(0)
    lbmcheck(ccall((:lbm_state_size, libhip), Cint, (Ptr{Cvoid}, Ref{Int64}), h, n))
    buf = Vector{UInt8}(undef, n[])
    lbmcheck(ccall((:lbm_save_state, libhip), Cint, (Ptr{Cvoid}, Ptr{UInt8}, Int64), h, buf, n[]))
    return buf
end
lbm_load_state(h, buf::Vector{UInt8}) =
    lbmcheck(ccall((:lbm_load_state, libhip), Cint, (Ptr{Cvoid}, Ptr{UInt8}, Int64), h, buf, length(buf)))
function lbm_check_populations(h)
    bad, sat = Ref{Int64}(0), Ref{Int64}(0)
    lbmcheck(ccall((:lbm_check_populations, libhip), Cint, (Ptr{Cvoid}, Ref{Int64}, Ref{Int64}), h, bad, sat))
    return bad[], sat[]
end
function lbm_autotune(h; reps = 20)
    best, ms = zeros(Int32, 3), Ref{Float64}(0.0)
    lbmcheck(ccall((:lbm_autotune, libhip), Cint, (Ptr{Cvoid}, Int32, Ptr{Int32}, Ref{Float64}), h, reps, best, ms))
    return Tuple(best), ms[]
end
lbm_viscous_stress!(h, sigma::Array{Float64}) =               # (N..., d, d)
    lbmcheck(ccall((:lbm_viscous_stress, libhip), Cint, (Ptr{Cvoid}, Ptr{Float64}), h, sigma))

# particle-particle interaction terms (molecularDynamicsTick!'s order)
lbm_set_interactions(h, terms::Vector{LbmInteraction}) =
    lbmcheck(ccall((:lbm_set_interactions, libhip), Cint, (Ptr{Cvoid}, Ptr{LbmInteraction}, Int32),
                   h, terms, length(terms)))

# per-tick particle log and the CSV writers of saveData
lbm_set_trajectory_log(h, capacity::Integer) =
    lbmcheck(ccall((:lbm_set_trajectory_log, libhip), Cint, (Ptr{Cvoid}, Int64), h, capacity))
function lbm_get_trajectory(h, id::Int, first::Integer, count::Integer)
    out = zeros(9, count)                                      # (position, velocity, angular velocity) per tick
    lbmcheck(ccall((:lbm_get_trajectory, libhip), Cint, (Ptr{Cvoid}, Int32, Int64, Int64, Ptr{Float64}),
                   h, id - 1, first, count, out))
    return out
end
function lbm_write_fluid_csv(path, dims, n3, tick, time, coords, rho, u, Q, f)
    cp = Ptr{Float64}[i <= dims ? pointer(coords[i]) : C_NULL for i in 1:3]
    GC.@preserve coords lbmcheck(ccall((:lbm_write_fluid_csv, libhip), Cint,
        (Cstring, Int32, Ptr{Int64}, Int64, Float64, Ptr{Ptr{Float64}}, Ptr{Float64}, Ptr{Float64}, Int32, Ptr{Float64}),
        path, dims, collect(Int64, n3), tick, time, cp, rho, u, Q, f))
end
function lbm_write_tensor_csv(path, dims, n3, tick, time, coords, t)
    cp = Ptr{Float64}[i <= dims ? pointer(coords[i]) : C_NULL for i in 1:3]
    GC.@preserve coords lbmcheck(ccall((:lbm_write_tensor_csv, libhip), Cint,
        (Cstring, Int32, Ptr{Int64}, Int64, Float64, Ptr{Ptr{Float64}}, Ptr{Float64}),
        path, dims, collect(Int64, n3), tick, time, cp, t))
end

# boundary links of a particle, (node linear index 0-based, direction 0-based)
function lbm_get_links(h, id::Int)
    cnt = Ref{Int64}(0)
    lbmcheck(ccall((:lbm_get_links, libhip), Cint, (Ptr{Cvoid}, Int32, Ptr{Int64}, Ptr{Int32}, Int64, Ref{Int64}),
                   h, id - 1, C_NULL, C_NULL, 0, cnt))
    node, dir = zeros(Int64, cnt[]), zeros(Int32, cnt[])
    lbmcheck(ccall((:lbm_get_links, libhip), Cint, (Ptr{Cvoid}, Int32, Ptr{Int64}, Ptr{Int32}, Int64, Ref{Int64}),
                   h, id - 1, node, dir, length(node), cnt))
    return node, dir
end

# batches (independent boxes advanced together)
function lbm_batch_create(hs::Vector{Ptr{Cvoid}})
    b = Ref{Ptr{Cvoid}}(C_NULL)
    lbmcheck(ccall((:lbm_batch_create, libhip), Cint, (Ptr{Ptr{Cvoid}}, Int32, Ref{Ptr{Cvoid}}), hs, length(hs), b))
    return b[]
end
lbm_batch_autotune(b; reps = 5) =
    lbmcheck(ccall((:lbm_batch_autotune, libhip), Cint, (Ptr{Cvoid}, Int32, Ptr{Int32}, Ptr{Float64}),
                   b, reps, C_NULL, C_NULL))
lbm_batch_run(b, n::Integer) = lbmcheck(ccall((:lbm_batch_run, libhip), Cint, (Ptr{Cvoid}, Int64), b, n))
lbm_batch_sync(b) = lbmcheck(ccall((:lbm_batch_sync, libhip), Cint, (Ptr{Cvoid},), b))
lbm_batch_destroy(b) = ccall((:lbm_batch_destroy, libhip), Cint, (Ptr{Cvoid},), b)

# the discrete adjoint of a swimming-speed loss
function lbm_adjoint_run(h, cfg::LbmAdjointConfig, np::Int, dF)
    loss = Ref{Float64}(0.0)
    dB1, dB2 = zeros(max(np, 1)), zeros(max(np, 1))
    lbmcheck(ccall((:lbm_adjoint_run, libhip), Cint,
                   (Ptr{Cvoid}, Ref{LbmAdjointConfig}, Ref{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                   h, cfg, loss, dB1, dB2, dF === nothing ? C_NULL : dF))
    return loss[], dB1[1:np], dB2[1:np]
end

# Julia's `range(a, stop = b, step = s)` values, as modelInit builds the axes
coordinateRange(lims, Δx) = collect(range(lims[1], stop = lims[2], step = Δx))

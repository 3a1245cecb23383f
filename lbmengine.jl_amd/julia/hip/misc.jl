#= hip backend of LBMengine.jl — misc.jl

Velocity sets, in the order of the reference's sets (gpu/misc.jl:195-272)
since the population index q is part of the public layout: rest velocity,
the axis velocities (-1, +1 per axis), then the face diagonals and (D3Q15,
D3Q27) the body diagonals, each group with its first component running
fastest as Julia's column-major comprehensions produce them. =#

# the four diagonals in the plane of axes (a, b), a fastest
planeDiagonals(a, b, d) = [begin
        v = zeros(Int8, d); v[a] = p; v[b] = q; v
    end for p in (-1, 1), q in (-1, 1)] |> vec

bodyDiagonals() = [Int8[p, q, r] for p in (-1, 1), q in (-1, 1), r in (-1, 1)] |> vec

axisVelocities(d) = vcat([zeros(Int8, d)], [begin
        v = zeros(Int8, d); v[k] = s; v
    end for s in (-1, 1), k in 1:d] |> vec)

function velocitySet(name::Symbol)
    if name == :D2Q9
        cs = vcat(axisVelocities(2), planeDiagonals(1, 2, 2))
        ws = vcat(4 / 9, fill(1 / 9, 4), fill(1 / 36, 4))
    elseif name == :D3Q15
        cs = vcat(axisVelocities(3), bodyDiagonals())
        ws = vcat(2 / 9, fill(1 / 9, 6), fill(1 / 72, 8))
    elseif name == :D3Q19
        cs = vcat(axisVelocities(3), planeDiagonals(2, 3, 3), planeDiagonals(1, 3, 3), planeDiagonals(1, 2, 3))
        ws = vcat(1 / 3, fill(1 / 18, 6), fill(1 / 36, 12))
    elseif name == :D3Q27
        cs = vcat(axisVelocities(3), planeDiagonals(2, 3, 3), planeDiagonals(1, 3, 3), planeDiagonals(1, 2, 3),
                  bodyDiagonals())
        ws = vcat(8 / 27, fill(2 / 27, 6), fill(1 / 54, 12), fill(1 / 216, 8))
    else
        error("velocity set $name is not available in the hip backend (:D2Q9, :D3Q15, :D3Q19 and :D3Q27 are)")
    end
    return (; cs, ws)
end

oppositeVectorId(cs) = [findfirst(c -> c == -ci, cs) for ci in cs]

# wallNodes (gpu/misc.jl:7-31): the first and last layer of every walled
# dimension, as a Bool array over the grid
function wallNodes(N::NTuple{D,Int}, walledDimensions) where {D}
    w = falses(N...)
    for k in walledDimensions
        idx = ntuple(j -> j == k ? 1 : Colon(), D)
        w[idx...] .= true
        idx = ntuple(j -> j == k ? N[k] : Colon(), D)
        w[idx...] .= true
    end
    return w
end

# the grid's node coordinates in Julia order (first index fastest)
gridPoints(coordinates) = Iterators.product(coordinates...)

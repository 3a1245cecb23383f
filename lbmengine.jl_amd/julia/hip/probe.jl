#= hip backend of LBMengine.jl — probe.jl

Included by src/LBMengine.jl before it picks the backend directory: the
library's path (LBM_HIP_LIB, else deps/liblbm_hip.so beside src/) and
whether it loads and sees a device. =#

const libhip = get(ENV, "LBM_HIP_LIB", joinpath(@__DIR__, "..", "..", "deps", "liblbm_hip.so"))

hipFunctional() = isfile(libhip) && begin
    n = Ref{Int32}(0)
    ccall((:lbm_device_count, libhip), Cint, (Ref{Int32},), n) == 0 && n[] > 0
end

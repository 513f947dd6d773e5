"""
    StokesFlowHIP

Drop-in replacements for the StokesFlow.jl hot path on one AMD MI355X (gfx950), calling
libsfhip through its C ABI (include/sf_hip.h).  The functions keep the reference's names,
argument order and return values:

  * `solveStokes!`          -- scripts/StokesSolver_multixpu.jl:14-96
  * `bilinearMarkerToGrid!` -- scripts/MarkerToGrid.jl:18-38
  * `moveMarkersRK4!`       -- scripts/MoveMarkers.jl:15-23
  * `saveStats`, `saveMarkers`, `saveGrid` -- scripts/StokesFlow2D_multixpu.jl:212-299
    (MAT-file level 5, which MAT.jl and viz.jl read)
  * `initializeMarkersCPU`, `setInitialMarkerProperties!` -- scripts/common.jl:40-110 (on the
    device; the density / viscosity callbacks become a `Materials` descriptor)

Arrays are `AMDGPU.ROCArray{Float64}` (column-major, x fastest: the reference's
`Data.Array` layout), passed to the library as device pointers.  One rank: `comm` and
`grid` are accepted and ignored, as ImplicitGlobalGrid/MPI are identities at one rank.

Julia is not available where this package is built and tested; the same ABI is exercised
by the Python binding (stokesflow.jl_amd/sfhip).
"""
module StokesFlowHIP

using AMDGPU

export solveStokes!, bilinearMarkerToGrid!, moveMarkersRK4!, saveStats, saveMarkers, saveGrid,
       initializeMarkersCPU, setInitialMarkerProperties!, Materials

const libsfhip = joinpath(@__DIR__, "..", "libsfhip.so")

const SF_OK = Cint(0)

struct SolveOut
    dt_new::Cdouble
    err::Cdouble
    t_eff::Cdouble
    runtime::Cdouble
    niter::Int64
    converged::Int32
    reserved::Int32
    dt_last::Cdouble
    t_it::Cdouble
    gbs_alg::Cdouble
end
SolveOut() = SolveOut(0, 0, 0, 0, 0, 0, 0, 0, 0, 0)

struct M2GOut
    n_bad_markers::Int64
    n_empty_nodes::Int64
end

struct RK4Out
    n_clamped::Int64
end

function check(rc::Cint)
    if rc != SF_OK
        msg = unsafe_string(ccall((:sf_last_error, libsfhip), Cstring, ()))
        error("libsfhip error $rc: $msg")
    end
    return nothing
end

# one context per (Nx, Ny, device); (0, 0) is the marker-only context
const CONTEXTS = Dict{Tuple{Int,Int,Int},Ptr{Cvoid}}()

function context(nx::Integer, ny::Integer)
    key = (Int(nx), Int(ny), AMDGPU.device_id(AMDGPU.device()))
    h = get(CONTEXTS, key, C_NULL)
    if h == C_NULL
        r = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:sf_ctx_create, libsfhip), Cint, (Ptr{Ptr{Cvoid}}, Cint, Cint, Ptr{Cvoid}),
                    r, nx, ny, C_NULL))
        h = r[]
        CONTEXTS[key] = h
    end
    # enqueue on the task-local HIP stream AMDGPU.jl uses
    check(ccall((:sf_ctx_set_stream, libsfhip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), h,
                AMDGPU.stream().stream))
    return h
end

dptr(A) = A === nothing ? Ptr{Float64}(C_NULL) : reinterpret(Ptr{Float64}, pointer(A))

"""
    solveStokes!(P, Vx_pad, Vy_pad, ρ_vy, μ_b, μ_p, τxx, τyy, τxy, ∇V, dτPt, Rx, Ry, dVxdτ, dVydτ,
                 dτVx, dτVy, Vx_s, Vy_s, g_y, dx, dy, Nx, Ny, dt, maxdisp, comm;
                 use_free_surface_stabilization=true, ϵ=1e-8, print_info=true, iterMax=100000)
        -> (dt, T_eff, runtime)
"""
function solveStokes!(P, Vx_pad, Vy_pad, ρ_vy, μ_b, μ_p, τxx, τyy, τxy, ∇V, dτPt, Rx, Ry, dVxdτ, dVydτ,
                      dτVx, dτVy, Vx_s, Vy_s, g_y, dx, dy, Nx, Ny, dt, maxdisp, comm;
                      use_free_surface_stabilization::Bool=true, ϵ=1e-8, print_info::Bool=true,
                      iterMax=100000)
    @assert size(P) == (Nx - 1, Ny - 1) && size(Vx_s) == (Nx, Ny + 1) && size(Vy_s) == (Nx + 1, Ny)
    @assert size(Vx_pad) == (Nx + 2, Ny + 1) && size(Vy_pad) == (Nx + 1, Ny + 2)
    h = context(Nx, Ny)
    out = Ref(SolveOut())
    check(ccall((:sf_solve_stokes, libsfhip), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
                 Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
                 Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
                 Ptr{Float64}, Ptr{Float64}, Cdouble, Cdouble, Cdouble, Cint, Cint, Cdouble, Cdouble,
                 Cint, Cdouble, Int64, Cint, Ptr{SolveOut}),
                h, dptr(P), dptr(Vx_pad), dptr(Vy_pad), dptr(ρ_vy), dptr(μ_b), dptr(μ_p), dptr(τxx),
                dptr(τyy), dptr(τxy), dptr(∇V), dptr(dτPt), dptr(Rx), dptr(Ry), dptr(dVxdτ), dptr(dVydτ),
                dptr(dτVx), dptr(dτVy), dptr(Vx_s), dptr(Vy_s), g_y, dx, dy, Nx, Ny, dt, maxdisp,
                use_free_surface_stabilization, ϵ, iterMax, print_info, out))
    return out[].dt_new, out[].t_eff, out[].runtime
end

"""
    bilinearMarkerToGrid!(x_grid_min, y_grid_min, val_grid, x_m, y_m, val_m, dx, dy, val_wt_sum, wt_sum, grid)
"""
function bilinearMarkerToGrid!(x_grid_min, y_grid_min, val_grid, x_m, y_m, val_m, dx, dy, val_wt_sum, wt_sum,
                               grid=nothing)
    Nx, Ny = size(val_grid)
    Nm = size(x_m, 1)
    @assert (size(y_m, 1) == Nm) && (size(val_m, 1) == Nm)
    h = context(0, 0)
    out = Ref(M2GOut(0, 0))
    wnx, wny = wt_sum === nothing ? (0, 0) : size(wt_sum)
    check(ccall((:sf_marker_to_grid, libsfhip), Cint,
                (Ptr{Cvoid}, Cdouble, Cdouble, Ptr{Float64}, Cint, Cint, Ptr{Float64}, Ptr{Float64},
                 Ptr{Float64}, Int64, Cdouble, Cdouble, Ptr{Float64}, Ptr{Float64}, Cint, Cint, Ptr{M2GOut}),
                h, x_grid_min, y_grid_min, dptr(val_grid), Nx, Ny, dptr(x_m), dptr(y_m), dptr(val_m), Nm,
                dx, dy, dptr(val_wt_sum), dptr(wt_sum), wnx, wny, out))
    out[].n_bad_markers > 0 && println("ATTENTION: erroneous marker position ($(out[].n_bad_markers) markers)")
    out[].n_empty_nodes > 0 &&
        println("PROBLEM: no markers close enough for value interpolation at $(out[].n_empty_nodes) nodes")
    return nothing
end

"""
    moveMarkersRK4!(x_m, y_m, Vx, Vy, x_vx_min, y_vx_min, x_vy_min, y_vy_min, dt, lx, ly, dx, dy)
"""
function moveMarkersRK4!(x_m, y_m, Vx, Vy, x_vx_min, y_vx_min, x_vy_min, y_vy_min, dt, lx, ly, dx, dy)
    Nm = length(x_m)
    @assert (Nm == length(y_m)) "x and y Marker coords have not same length"
    h = context(0, 0)
    out = Ref(RK4Out(0))
    check(ccall((:sf_move_markers_rk4, libsfhip), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Int64, Ptr{Float64}, Cint, Cint, Ptr{Float64}, Cint,
                 Cint, Cdouble, Cdouble, Cdouble, Cdouble, Cdouble, Cdouble, Cdouble, Cdouble, Cdouble,
                 Ptr{RK4Out}),
                h, dptr(x_m), dptr(y_m), Nm, dptr(Vx), size(Vx, 1), size(Vx, 2), dptr(Vy), size(Vy, 1),
                size(Vy, 2), x_vx_min, y_vx_min, x_vy_min, y_vy_min, dt, lx, ly, dx, dy, out))
    return nothing
end

"""
    saveStats(nt, rank, dims)   -- StokesFlow2D_multixpu.jl:212-223 (writes viz_out/stats.mat on rank 0)
"""
function saveStats(nt, rank, dims)
    (rank != 0) && return nothing
    d = Int64[dims[1], dims[2], dims[3]]
    check(ccall((:sf_save_stats, libsfhip), Cint, (Cstring, Ptr{Int64}, Int64), "viz_out/stats.mat", d, nt))
    return nothing
end

"""
    saveMarkers(nt, rank, coords, localDomain, dx, dy, x_m, y_m, ρ_m)   -- :230-240
"""
function saveMarkers(nt, rank, coords, localDomain, dx, dy, x_m, y_m, ρ_m)
    path = string("viz_out/markers_", lpad(nt, 4, '0'), "_", lpad(rank, 4, '0'), ".mat")
    h = context(0, 0)
    check(ccall((:sf_save_markers, libsfhip), Cint,
                (Ptr{Cvoid}, Cstring, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Int64, Cdouble, Cdouble),
                h, path, dptr(x_m), dptr(y_m), dptr(ρ_m), length(x_m), (localDomain[1] - dx) * coords[1],
                (localDomain[2] - dy) * coords[2]))
    return nothing
end

"""
    saveGrid(nt, rank, x, y, μ_b, x_p, y_p, P, x_vx, y_vx, Vx, x_vy, y_vy, Vy)   -- :247-299

The coordinate vectors are recomputed by the library from the grid step (x[2] - x[1]).
"""
function saveGrid(nt, rank, x, y, μ_b, x_p, y_p, P, x_vx, y_vx, Vx, x_vy, y_vy, Vy)
    path = string("viz_out/grid_", lpad(nt, 4, '0'), "_", lpad(rank, 4, '0'), ".mat")
    h = context(0, 0)
    check(ccall((:sf_save_grid, libsfhip), Cint,
                (Ptr{Cvoid}, Cstring, Cint, Cint, Cdouble, Cdouble, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
                 Ptr{Float64}),
                h, path, length(x), length(y), x[2] - x[1], y[2] - y[1], dptr(μ_b), dptr(P), dptr(Vx), dptr(Vy)))
    return nothing
end

"""
    initializeMarkersCPU(comm, dims, coords, marker_density, lx, ly, dx, dy, Nx, Ny, RAND_MARKER_POS; seed=0)
        -- common.jl:40-100 at one rank

The marker lattice, built on the device bit for bit as the reference builds it; returns
(x_m, y_m, ρ_m, μ_m) as ROCArrays (ρ_m, μ_m zero).  With RAND_MARKER_POS the jitter comes from
the library's seeded generator (`seed`), not from Julia's RNG.
"""
function initializeMarkersCPU(comm, dims, coords, marker_density::Integer, lx, ly, dx, dy, Nx, Ny, RAND_MARKER_POS;
                              seed=0)
    nm = marker_density^2 * (Nx - 1) * (Ny - 1)
    x_m = AMDGPU.zeros(Float64, nm)
    y_m = AMDGPU.zeros(Float64, nm)
    h = context(0, 0)
    check(ccall((:sf_lattice_markers, libsfhip), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Cint, Cint, Cint, Cdouble, Cdouble, Cdouble, Cdouble, Cint,
                 UInt64),
                h, dptr(x_m), dptr(y_m), Nx, Ny, marker_density, lx, ly, dx, dy, RAND_MARKER_POS ? 1 : 0, seed))
    return x_m, y_m, AMDGPU.zeros(Float64, nm), AMDGPU.zeros(Float64, nm)
end

struct Shape
    kind::Int32
    reserved::Int32
    a::Cdouble
    b::Cdouble
    c::Cdouble
    d::Cdouble
    e::Cdouble
    ρ::Cdouble
    μ::Cdouble
end

const SHAPE_KINDS = Dict(:y_below => 1, :y_above => 2, :circle => 3, :rect => 4, :sine_below => 5)

"""
    Materials(shapes, ρ_bg, μ_bg)

Material descriptor in place of the `density(x, y)` / `viscosity(x, y)` callbacks: shapes in
order, `(:y_below, h, ρ, μ)`, `(:y_above, h, ρ, μ)`, `(:circle, cx, cy, r, ρ, μ)`,
`(:rect, x0, x1, y0, y1, ρ, μ)`, `(:sine_below, y0, amp, k, lx, ρ, μ)`; the first shape that
contains a point decides, else the background.  The callbacks of StokesFlow2D_multixpu.jl:387-404
are `Materials([(:y_below, air_height, ρ_air, μ_air), (:circle, plume_x, plume_y, plume_r,
ρ_plume, μ_plume)], ρ_matrix, μ_matrix)`.
"""
struct Materials
    shapes::Vector{Shape}
    ρ_bg::Float64
    μ_bg::Float64
end
function Materials(shapes::AbstractVector, ρ_bg, μ_bg)
    v = Shape[]
    for sh in shapes
        kind, vals... = sh
        p = collect(Float64, vals[1:end-2])
        append!(p, zeros(5 - length(p)))
        push!(v, Shape(SHAPE_KINDS[kind], 0, p[1], p[2], p[3], p[4], p[5], vals[end-1], vals[end]))
    end
    return Materials(v, ρ_bg, μ_bg)
end

"""
    setInitialMarkerProperties!(x_m, y_m, ρ_m, μ_m, x0, y0, materials::Materials)   -- common.jl:104-110
"""
function setInitialMarkerProperties!(x_m, y_m, ρ_m, μ_m, x0, y0, materials::Materials)
    Nm = length(x_m)
    @assert (length(y_m) == Nm) && (length(ρ_m) == Nm) && (length(μ_m) == Nm)
    h = context(0, 0)
    check(ccall((:sf_marker_properties, libsfhip), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Int64, Cdouble, Cdouble, Ptr{Shape}, Cint, Cdouble, Cdouble,
                 Ptr{Float64}, Ptr{Float64}),
                h, dptr(x_m), dptr(y_m), Nm, x0, y0, materials.shapes, length(materials.shapes), materials.ρ_bg,
                materials.μ_bg, dptr(ρ_m), dptr(μ_m)))
    return nothing
end

end # module

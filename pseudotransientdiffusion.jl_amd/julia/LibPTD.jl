# LibPTD.jl -- Julia binding of libptd (include/ptd.h), the gfx950 HIP library that replaces the
# reference's ParallelStencil kernels and host PT loop (scripts/diff_3D/diff_3D_nonlin_multixpu.jl:37-55,
# 99-119; scripts/diff_3D/diff_3D_nonlin_perf.jl:32-59,93-112 and the 1D / 2D scripts' equivalents).
#
# Arrays stay Julia `Array{Float64}`s in the scripts' own shapes (column-major, the staggered flux shapes
# of scripts/diff_3D/diff_3D_nonlin_multixpu.jl:77-83); the library owns the device copies inside a
# context. Every call is synchronous. A status other than PTD_OK raises; PTD_ERR_NAN raises "NaN" as the
# scripts' `isnan(err) && error("NaN")` does (scripts/diff_3D/diff_3D_nonlin_multixpu.jl:118).
module LibPTD

export PtdParams, Context, create, destroy, set_params!, set!, get!, field_count, device_bytes,
       compute_iter_params!, compute_flux!, compute_update!, compute_flux_res!, check_res!, assign!,
       iterate!, iterate_timed!, iterations_per_launch, pt_solve!, solve!, outer_iterations, grad_solve!,
       Batch, batch_create, batch_destroy, batch_set!, batch_get!, batch_solve!, batch_tiling,
       batch_group_times

const libptd = get(ENV, "LIBPTD", joinpath(@__DIR__, "..", "libptd.so"))

# ptd_status
const PTD_OK, PTD_ERR_ARG, PTD_ERR_HIP, PTD_ERR_NAN, PTD_ERR_NOMEM, PTD_ERR_STATE, PTD_ERR_NODEV = 0, 1, 2, 3, 4, 5, 6
# ptd_dcase, ptd_form, ptd_precision, ptd_arith, ptd_norm
const PTD_LIN, PTD_STEP, PTD_NONLIN = 0, 1, 2
const PTD_PLAIN, PTD_PERF = 0, 1
const PTD_FP64, PTD_FP32, PTD_MIXED = 0, 1, 2
const PTD_ARITH_EXACT, PTD_ARITH_FAST = 0, 1
const PTD_NORM_RMS, PTD_NORM_MAX = 0, 1
# ptd_field
const PTD_F_H, PTD_F_HOLD, PTD_F_D, PTD_F_QX, PTD_F_QY, PTD_F_QZ = 0, 1, 2, 3, 4, 5
const PTD_F_QX2, PTD_F_QY2, PTD_F_QZ2, PTD_F_RESH, PTD_F_DTAU = 6, 7, 8, 9, 10

# include/ptd.h: ptd_params (4 x Int32, 4 x Int64, 4 x Int32, 14 x Float64 = 176 bytes, natural alignment)
struct PtdParams
    ndim::Int32
    dcase::Int32
    form::Int32
    precision::Int32
    nx::Int64
    ny::Int64
    nz::Int64
    itmax::Int64
    nout::Int32
    stop_norm::Int32
    arith::Int32
    reserved::Int32
    dx::Float64
    dy::Float64
    dz::Float64
    dt::Float64
    vpdt::Float64
    max_l::Float64
    max_l2::Float64
    resc::Float64
    d0::Float64
    re::Float64
    theta_r::Float64
    dtau_rho::Float64
    tol::Float64
    norm_len::Float64
end

"""
    PtdParams(ndim, dcase, form, n, itMax, nout, dx, dt, Vpdτ, max_l, Resc, tol; kwargs...)

The scalar block a script derives in its "Derived numerics" section (e.g.
scripts/diff_3D/diff_3D_nonlin_multixpu.jl:70-74, scripts/diff_3D/diff_3D_lin_multixpu.jl:58-62). `n` and
`dx` are tuples over the active axes; lin / step scripts pass `d0`, `re`, `theta_r`, `dtau_rho`.
"""
function PtdParams(ndim, dcase, form, n, itMax, nout, dx, dt, Vpdτ, max_l, Resc, tol;
                   precision=PTD_FP64, arith=PTD_ARITH_EXACT, stop_norm=PTD_NORM_RMS, d0=1.0, re=0.0,
                   theta_r=0.0, dtau_rho=0.0, norm_len=0.0)
    nn = (n..., 1, 1)
    hh = (dx..., 1.0, 1.0)
    return PtdParams(Int32(ndim), Int32(dcase), Int32(form), Int32(precision),
                     Int64(nn[1]), Int64(nn[2]), Int64(nn[3]), Int64(itMax),
                     Int32(nout), Int32(stop_norm), Int32(arith), Int32(0),
                     Float64(hh[1]), Float64(hh[2]), Float64(hh[3]), Float64(dt), Float64(Vpdτ),
                     Float64(max_l), Float64(max_l)^2, Float64(Resc), Float64(d0), Float64(re),
                     Float64(theta_r), Float64(dtau_rho), Float64(tol), Float64(norm_len))
end

last_error() = unsafe_string(ccall((:ptd_last_error, libptd), Cstring, ()))

function check(st::Integer)
    st == PTD_OK && return nothing
    st == PTD_ERR_NAN && error("NaN")
    error("libptd status $(st): ", last_error())
end

abi_version() = ccall((:ptd_abi_version, libptd), Int32, ())

function device_check(dev::Integer=0)
    name = Vector{UInt8}(undef, 64)
    st = ccall((:ptd_device_check, libptd), Int32, (Int32, Ptr{UInt8}, Int32), dev, name, length(name))
    return st, unsafe_string(pointer(name))
end

# ---- contexts ------------------------------------------------------------------------------------

mutable struct Context
    handle::Ptr{Cvoid}
    params::PtdParams
end

function create(p::PtdParams; device::Integer=0)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:ptd_ctx_create, libptd), Int32, (Ref{PtdParams}, Int32, Ref{Ptr{Cvoid}}), p, device, h))
    ctx = Context(h[], p)
    finalizer(destroy, ctx)
    return ctx
end

function destroy(ctx::Context)
    if ctx.handle != C_NULL
        ccall((:ptd_ctx_destroy, libptd), Int32, (Ptr{Cvoid},), ctx.handle)
        ctx.handle = C_NULL
    end
    return nothing
end

function set_params!(ctx::Context, p::PtdParams)
    check(ccall((:ptd_ctx_set_params, libptd), Int32, (Ptr{Cvoid}, Ref{PtdParams}), ctx.handle, p))
    ctx.params = p
    return ctx
end

device_bytes(ctx::Context) = ccall((:ptd_ctx_device_bytes, libptd), Int64, (Ptr{Cvoid},), ctx.handle)
outer_iterations(ctx::Context) = ccall((:ptd_ctx_outer_iterations, libptd), Int64, (Ptr{Cvoid},), ctx.handle)
stream(ctx::Context) = ccall((:ptd_ctx_stream, libptd), Ptr{Cvoid}, (Ptr{Cvoid},), ctx.handle)
field_count(ctx::Context, f::Integer) = ccall((:ptd_field_count, libptd), Int64, (Ptr{Cvoid}, Int32), ctx.handle, f)

set!(ctx::Context, f::Integer, A::Array{Float64}) =
    check(ccall((:ptd_set_field, libptd), Int32, (Ptr{Cvoid}, Int32, Ptr{Float64}, Int64), ctx.handle, f, A, length(A)))

function get!(A::Array{Float64}, ctx::Context, f::Integer)
    check(ccall((:ptd_get_field, libptd), Int32, (Ptr{Cvoid}, Int32, Ptr{Float64}, Int64), ctx.handle, f, A, length(A)))
    return A
end

# ---- kernel level: one reference kernel launch each ---------------------------------------------

compute_iter_params!(ctx::Context) = check(ccall((:ptd_compute_iter_params, libptd), Int32, (Ptr{Cvoid},), ctx.handle))
compute_flux!(ctx::Context)        = check(ccall((:ptd_compute_flux, libptd), Int32, (Ptr{Cvoid},), ctx.handle))
compute_update!(ctx::Context)      = check(ccall((:ptd_compute_update, libptd), Int32, (Ptr{Cvoid},), ctx.handle))
compute_flux_res!(ctx::Context)    = check(ccall((:ptd_compute_flux_res, libptd), Int32, (Ptr{Cvoid},), ctx.handle))
assign!(ctx::Context)              = check(ccall((:ptd_assign, libptd), Int32, (Ptr{Cvoid},), ctx.handle))

function check_res!(ctx::Context)
    rms, mx = Ref(0.0), Ref(0.0)
    check(ccall((:ptd_check_res, libptd), Int32, (Ptr{Cvoid}, Ref{Float64}, Ref{Float64}), ctx.handle, rms, mx))
    return rms[], mx[]
end

# ---- loop level ---------------------------------------------------------------------------------

iterate!(ctx::Context, niter::Integer) = check(ccall((:ptd_iterate, libptd), Int32, (Ptr{Cvoid}, Int64), ctx.handle, niter))

function iterate_timed!(ctx::Context, niter::Integer)
    ms = Ref(0.0)
    check(ccall((:ptd_iterate_timed, libptd), Int32, (Ptr{Cvoid}, Int64, Ref{Float64}), ctx.handle, niter, ms))
    return ms[]
end

# PT iterations per launch of the context's unchecked iterations (3 three-level, 2 two-level, 1 single sweeps)
function iterations_per_launch(ctx::Context)
    n = Ref(Int32(0))
    check(ccall((:ptd_ctx_iterations_per_launch, libptd), Int32, (Ptr{Cvoid}, Ref{Int32}), ctx.handle, n))
    return Int(n[])
end

function pt_solve!(ctx::Context)
    it, rms, mx = Ref(Int64(0)), Ref(0.0), Ref(0.0)
    check(ccall((:ptd_pt_solve, libptd), Int32, (Ptr{Cvoid}, Ref{Int64}, Ref{Float64}, Ref{Float64}),
                ctx.handle, it, rms, mx))
    return it[], rms[], mx[]
end

function solve!(ctx::Context, nt::Integer)
    iters, errs = zeros(Int64, nt), zeros(Float64, nt)
    check(ccall((:ptd_solve, libptd), Int32, (Ptr{Cvoid}, Int32, Ptr{Int64}, Ptr{Float64}),
                ctx.handle, nt, iters, errs))
    return iters, errs
end

"""
    grad_solve!(ctx, nt, Hobs) -> (L, dL/dD, dL/dH0, iters_fwd, iters_adj)

nt physical steps of the context's problem and the gradient of L = 1/2 sum_inner (H(T) - H_obs)^2 by the
adjoint (no reference counterpart); dL/dD is `nothing` for D = H^3. The context ends as after
`solve!(ctx, nt)`.
"""
function grad_solve!(ctx::Context, nt::Integer, Hobs::Array{Float64})
    L = Ref(0.0)
    nonlin = ctx.params.dcase == PTD_NONLIN
    gD, gH0 = nonlin ? nothing : similar(Hobs), similar(Hobs)
    itf, ita = zeros(Int64, nt), zeros(Int64, nt)
    check(ccall((:ptd_grad_solve, libptd), Int32,
                (Ptr{Cvoid}, Int32, Ptr{Float64}, Int64, Ref{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Int64}, Ptr{Int64}),
                ctx.handle, nt, Hobs, length(Hobs), L, nonlin ? Ptr{Float64}(C_NULL) : pointer(gD), gH0, itf, ita))
    return L[], gD, gH0, itf, ita
end

# ---- batches of independent solves (no reference counterpart) -----------------------------------

mutable struct Batch
    handle::Ptr{Cvoid}
    n::Int
end

function batch_create(ps::Vector{PtdParams}; device::Integer=0)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:ptd_batch_create, libptd), Int32, (Ptr{PtdParams}, Int32, Int32, Ref{Ptr{Cvoid}}),
                ps, length(ps), device, h))
    b = Batch(h[], length(ps))
    finalizer(batch_destroy, b)
    return b
end

function batch_destroy(b::Batch)
    if b.handle != C_NULL
        ccall((:ptd_batch_destroy, libptd), Int32, (Ptr{Cvoid},), b.handle)
        b.handle = C_NULL
    end
    return nothing
end

batch_size(b::Batch) = ccall((:ptd_batch_size, libptd), Int32, (Ptr{Cvoid},), b.handle)
batch_device_bytes(b::Batch) = ccall((:ptd_batch_device_bytes, libptd), Int64, (Ptr{Cvoid},), b.handle)

# members are numbered from 1 here, from 0 in the C ABI
batch_set!(b::Batch, m::Integer, f::Integer, A::Array{Float64}) =
    check(ccall((:ptd_batch_set_field, libptd), Int32, (Ptr{Cvoid}, Int32, Int32, Ptr{Float64}, Int64),
                b.handle, m - 1, f, A, length(A)))

function batch_get!(A::Array{Float64}, b::Batch, m::Integer, f::Integer)
    check(ccall((:ptd_batch_get_field, libptd), Int32, (Ptr{Cvoid}, Int32, Int32, Ptr{Float64}, Int64),
                b.handle, m - 1, f, A, length(A)))
    return A
end

"""
    batch_solve!(b, nt) -> (iters[n, nt], errs[n, nt], latency_ms[n], status[n], total_ms)

nt physical steps of every member together; each member's result equals `solve!` on it alone.
"""
function batch_solve!(b::Batch, nt::Integer)
    iters = zeros(Int64, nt, b.n)      # the library writes member-major: column m is member m
    errs = zeros(Float64, nt, b.n)
    lat = zeros(Float64, b.n)
    status = zeros(Int32, b.n)
    total = Ref(0.0)
    st = ccall((:ptd_batch_solve, libptd), Int32,
               (Ptr{Cvoid}, Int32, Ptr{Int64}, Ptr{Float64}, Ptr{Float64}, Ptr{Int32}, Ref{Float64}),
               b.handle, nt, iters, errs, lat, status, total)
    st == PTD_ERR_NAN || check(st)
    return permutedims(iters), permutedims(errs), lat, status, total[]
end

function batch_tiling(b::Batch)
    tz, rows = Ref(Int32(0)), Ref(Int32(0))
    ms = zeros(Float64, 8)
    check(ccall((:ptd_batch_tiling, libptd), Int32, (Ptr{Cvoid}, Ref{Int32}, Ref{Int32}, Ptr{Float64}, Int32),
                b.handle, tz, rows, ms, length(ms)))
    return tz[], rows[], ms
end

"""
    batch_group_times(b) -> Vector of (kind, dcase, members, ms)

Per member group of the last `batch_solve!`: sweep kind (3: 3D two-level, 22: 2D two-level, 2: 2D), D case,
member count and the device time the group's stream took, summed over the replays.
"""
function batch_group_times(b::Batch; cap::Integer=64)
    ng = Ref(Int32(0))
    kind, dcase, mem = zeros(Int32, cap), zeros(Int32, cap), zeros(Int32, cap)
    ms = zeros(Float64, cap)
    check(ccall((:ptd_batch_group_times, libptd), Int32,
                (Ptr{Cvoid}, Ref{Int32}, Ptr{Int32}, Ptr{Int32}, Ptr{Int32}, Ptr{Float64}, Int32),
                b.handle, ng, kind, dcase, mem, ms, cap))
    n = min(Int(ng[]), cap)
    return [(kind[g], dcase[g], mem[g], ms[g]) for g in 1:n]
end

end # module

# diff_3D_nonlin_multixpu.jl on libptd: scripts/diff_3D/diff_3D_nonlin_multixpu.jl of the reference with
# its ParallelStencil kernels and PT loop replaced by LibPTD calls, on one rank (one GPU). The environment
# interface, physics, numerics, derived scalars, initial condition, printed line and DO_SAVE row are the
# script's (scripts/diff_3D/diff_3D_nonlin_multixpu.jl:1-8,57-89,120-133). On one rank ImplicitGlobalGrid's
# x_g(ix) is (ix-1)*dx and len_ResH_g is (nx-2)(ny-2)(nz-2).
#
#   USE_RETURN=true NX=32 NY=32 NZ=32 julia -O3 --check-bounds=no diff_3D_nonlin_multixpu.jl
const use_return  = haskey(ENV, "USE_RETURN" ) ? parse(Bool, ENV["USE_RETURN"] ) : false
const do_save     = haskey(ENV, "DO_SAVE"    ) ? parse(Bool, ENV["DO_SAVE"]    ) : false
const nx          = haskey(ENV, "NX"         ) ? parse(Int , ENV["NX"]         ) : 64
const ny          = haskey(ENV, "NY"         ) ? parse(Int , ENV["NY"]         ) : 64
const nz          = haskey(ENV, "NZ"         ) ? parse(Int , ENV["NZ"]         ) : 64
const arith       = haskey(ENV, "PTD_ARITH"  ) ? parse(Int , ENV["PTD_ARITH"]  ) : 0   # 0 EXACT, 1 FAST
using Printf
include(joinpath(@__DIR__, "LibPTD.jl"))
using .LibPTD

function diffusion_3D_()
    # Physics
    lx, ly, lz = 10.0, 10.0, 10.0 # domain size
    ttot       = 1.0              # total simulation time
    dt         = 0.2              # physical time step
    # Numerics
    tol        = 1e-8             # tolerance
    itMax      = 1e5              # max number of iterations
    nout       = 10               # tol check
    CFL        = 1 / sqrt(3)      # CFL number
    Resc       = 1 / 1.2          # iteration parameter scaling
    # Derived numerics (one rank: nx_g() = nx)
    dx, dy, dz = lx/nx, ly/ny, lz/nz
    Vpdτ       = CFL * min(dx, dy, dz)
    max_lxyz   = max(lx, ly, lz)
    xc, yc, zc = LinRange(dx/2, lx-dx/2, nx), LinRange(dy/2, ly-dy/2, ny), LinRange(dz/2, lz-dz/2, nz)
    # Initial condition, x_g(ix) = (ix-1)*dx on one rank
    H0         = [exp(-((ix-1)*dx-0.5*lx+dx/2)*((ix-1)*dx-0.5*lx+dx/2) - ((iy-1)*dy-0.5*ly+dy/2)*((iy-1)*dy-0.5*ly+dy/2) -
                      ((iz-1)*dz-0.5*lz+dz/2)*((iz-1)*dz-0.5*lz+dz/2)) for ix=1:nx, iy=1:ny, iz=1:nz]
    len_ResH_g = (nx-2)*(ny-2)*(nz-2)
    nt         = Int(ceil(ttot/dt))
    # The PT loop on the GPU: nt x {while err > tol && iter < itMax ...; Hold .= H} (:99-119)
    p   = LibPTD.PtdParams(3, LibPTD.PTD_NONLIN, LibPTD.PTD_PLAIN, (nx, ny, nz), itMax, nout, (dx, dy, dz), dt,
                           Vpdτ, max_lxyz, Resc, tol; arith=arith, norm_len=len_ResH_g)
    ctx = LibPTD.create(p)
    LibPTD.set!(ctx, LibPTD.PTD_F_H, H0)
    LibPTD.set!(ctx, LibPTD.PTD_F_HOLD, H0)
    iters, errs = LibPTD.solve!(ctx, nt)      # raises "NaN" as the script does
    ittot = sum(iters)
    H = LibPTD.get!(similar(H0), ctx, LibPTD.PTD_F_H)
    LibPTD.destroy(ctx)
    @printf("Total time = %1.2f, time steps = %d, nx = %d, iterations tot = %d \n", round(ttot, sigdigits=2), nt, nx, ittot)
    if do_save
        !ispath("../output") && mkdir("../output")
        open("../output/out_diff_3D_nonlin.txt","a") do io
            println(io, "$(nx) $(ny) $(nz) $(ittot) $(nt)")
        end
    end
    return xc, yc, zc, H
end

if use_return
    xc, yc, zc, H = diffusion_3D_();
else
    diffusion_3D = begin diffusion_3D_(); return; end
end

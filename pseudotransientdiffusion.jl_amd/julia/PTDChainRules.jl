# PTDChainRules.jl -- a ChainRulesCore rrule for the loss of an nt-step solve on libptd (lin, step, D = H^3)
# (no reference counterpart: the reference has no derivative). The pullback scales the adjoint gradients
# ptd_grad_solve returns (dL/dD, dL/dH0) by the incoming cotangent.
#
#   include("LibPTD.jl"); include("PTDChainRules.jl")
#   L, back = ChainRulesCore.rrule(PTDChainRules.solve_loss, ctx, D, H0, Hobs, nt)
module PTDChainRules

using ChainRulesCore
using ..LibPTD

inner(A::Array{Float64,3}) = view(A, 2:size(A,1)-1, 2:size(A,2)-1, 2:size(A,3)-1)
inner(A::Array{Float64,2}) = view(A, 2:size(A,1)-1, 2:size(A,2)-1)

function prepare!(ctx, D, H0)
    ctx.params.dcase == LibPTD.PTD_STEP && LibPTD.set!(ctx, LibPTD.PTD_F_D, D)
    LibPTD.set!(ctx, LibPTD.PTD_F_H, H0)
    LibPTD.set!(ctx, LibPTD.PTD_F_HOLD, H0)
end

"""
    solve_loss(ctx, D, H0, Hobs, nt) = 1/2 sum_inner (H(T) - Hobs)^2

H(T) after nt backward-Euler steps of the context's problem from H0 (with D the step field).
"""
function solve_loss(ctx, D, H0, Hobs, nt)
    prepare!(ctx, D, H0)
    LibPTD.solve!(ctx, nt)
    H = LibPTD.get!(similar(H0), ctx, LibPTD.PTD_F_H)
    return 0.5 * sum(abs2, inner(H) .- inner(Hobs))
end

function ChainRulesCore.rrule(::typeof(solve_loss), ctx, D, H0, Hobs, nt)
    prepare!(ctx, D, H0)
    L, gD, gH0, _, _ = LibPTD.grad_solve!(ctx, nt, Hobs)
    dD(ΔL) = gD === nothing ? NoTangent() : ΔL .* gD   # D = H^3: no D parameter
    pullback(ΔL) = (NoTangent(), NoTangent(), dD(ΔL), ΔL .* gH0, @not_implemented("dL/dHobs"), NoTangent())
    return L, pullback
end

end # module

# ChainRulesCore rrule of FFHIP.solve: the discrete adjoint of the fused
# ETDRK4 run (libffhip's ffhip_op_etdrk4_adjoint, checkpoint and recompute).
# No counterpart in FourierFlows.jl, which has no AD dependency.
#
# The reverse sweep needs the checkpoints of this very forward run.  A later
# solve with the same operator replaces them (the operator holds one set), so
# the pullback compares the operator's checkpoint run id with the one the
# forward pass recorded and, if they differ, runs the checkpointed forward
# again from a saved copy of sol0 before the adjoint.
#
# Complex cotangents: Julia's convention dL/dRe + i dL/dIm is the transpose for
# the real inner product Re sum conj(a) b that libffhip's adjoint uses, so the
# cotangent of solT goes in as is and the result is the cotangent of sol0.
module FFHIPChainRulesCoreExt

using ChainRulesCore
using FFHIP: FFHIP, BuiltinOperator, ETDRK4TimeStepper, HIPArray, solve, checkpoint_id,
             etdrk4_forward_ckpt!, etdrk4_adjoint!

function ChainRulesCore.rrule(::typeof(solve), op::BuiltinOperator, ts::ETDRK4TimeStepper, sol0::HIPArray,
                              nsteps::Integer, ncheck::Integer)
    start = copy(sol0)  # sol0 may be overwritten by the caller before the pullback runs
    solT = solve(op, ts, sol0, nsteps, ncheck)
    run = checkpoint_id(op)
    function solve_pullback(dsolT)
        lam = unthunk(dsolT)
        lam isa AbstractZero && return (NoTangent(), NoTangent(), NoTangent(), ZeroTangent(), NoTangent(),
                                        NoTangent())
        if checkpoint_id(op) != run  # another solve replaced the checkpoints: rebuild them
            etdrk4_forward_ckpt!(op, ts, similar(start), start, nsteps, ncheck)
        end
        lam0 = copy(lam::HIPArray)
        etdrk4_adjoint!(op, ts, lam0, lam0, nsteps, ncheck)  # in place: lam0 <- (dsolT/dsol0)^T lam
        return (NoTangent(), NoTangent(), NoTangent(), lam0, NoTangent(), NoTangent())
    end
    return solT, solve_pullback
end

end # module

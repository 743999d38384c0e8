"""A user boundary that writes interior cells sees the reference's order.

MISO applies the boundary after every RK stage and every AV sweep
(mhd_integrator_host.hpp:250-300) and reduces the CFL from the state the last
one leaves (mhd_model_base.hpp:57-62).  A boundary functor may write interior
cells -- density floors, blending near an obstacle
(demo/mhd3d_magnetosphere/src/bc.hpp:125-157) -- so with a user boundary set
the library must not fold a sweep, the CFL reduction or the radiative heating
into an earlier pass.  Each test composes the step from the per-function
entry points (misohip_cfl, misohip_rk_stage, misohip_ghost_fill,
misohip_av_sweep) in the reference's order and checks misohip_run,
misohip_step and misohip_update_heated against it bit for bit, dts included,
with a floor functor that clips a third of the cells on every call.
"""
import ctypes as C

import numpy as np
import pytest

GM = 5.0 / 3.0
CASES = {"3d": (20, 16, 12), "2d": (28, 24, 1)}


def _setup(ctx, sizes):
    from miso_amd import mhd, problems
    g = mhd.Grid(*sizes, 2, device="cuda:0")
    if sizes[2] > 1:
        q0 = problems.turbulence(g, device="cpu", seed=7)
    else:
        q0 = problems.orszag_tang(g, device="cpu")
    q0 = problems.to_numpy(q0)
    # floors at the 35th percentile of the initial density and internal
    # energy: the functor clips cells inside the domain on every call
    ro_floor = float(np.quantile(q0[0], 0.35))
    ei_floor = float(np.quantile(q0[7], 0.35))

    def floor(t):
        t[0].clamp_(min=ro_floor)
        t[7].clamp_(min=ei_floor)

    bc = mhd.BoundaryCondition((True, True, True), user=floor)
    m = mhd.MHD(mhd.Config({"mhd": {"cfl_number": 0.4}}), g, ctx, GM)
    m.apply_initial_condition(q0, bc)
    return m, bc, q0


def _reference_order_step(m, bc):
    """One ModelBase::update from the per-function entry points, in the
    reference's order; returns dt."""
    integ, qq = m.integrator, m.qq
    dt = integ.cfl(qq)
    ch2, tau = integ.divb_parameters(dt)
    argm = qq
    outs = (integ.qq_rslt, integ.qq_argm, integ.qq_rslt, integ.qq_argm)
    for s, frac in enumerate((4.0, 3.0, 2.0, 1.0)):
        out = outs[s]
        integ.update_sc4(dt / frac, float(tau), float(ch2), qq, argm, out)
        integ.apply_boundary_condition(bc, out)
        argm = out
    cur, other = argm, integ.qq_argm if argm is integ.qq_rslt else integ.qq_rslt
    integ.av_sweep(0, dt, cur, other, integ.cc)
    integ.apply_boundary_condition(bc, other)
    integ.av_sweep(1, dt, other, cur, integ.cc)
    integ.apply_boundary_condition(bc, cur)
    integ.av_sweep(2, dt, cur, qq, integ.cc)
    integ.apply_boundary_condition(bc, qq)
    return dt


@pytest.mark.gpu
@pytest.mark.parametrize("case", list(CASES))
def test_run_and_step_follow_the_reference_order(ctx, case):
    import torch
    nsteps = 4
    results = []
    for mode in ("reference", "run", "step"):
        m, bc, q0 = _setup(ctx, CASES[case])
        if mode == "reference":
            dts = [_reference_order_step(m, bc) for _ in range(nsteps)]
        elif mode == "run":
            dts = m.integrator.run(bc, m.qq, nsteps)
        else:
            dts = [m.integrator.step(bc, m.qq) for _ in range(nsteps)]
        torch.cuda.synchronize()
        results.append((mode, dts, m.qq.data.clone()))
    ref_dts, ref_q = results[0][1], results[0][2]
    assert all(d > 0 for d in ref_dts)
    for mode, dts, q in results[1:]:
        assert dts == ref_dts, (case, mode, dts, ref_dts)
        assert torch.equal(q, ref_q), (case, mode)
    # the functor mattered: without it the same run ends elsewhere
    from miso_amd import mhd
    m, _, q0 = _setup(ctx, CASES[case])
    plain = mhd.BoundaryCondition((True, True, True))
    m.qq.copy_from_host(ctx, q0)
    m.integrator.apply_boundary_condition(plain, m.qq)
    dts = m.integrator.run(plain, m.qq, nsteps)
    torch.cuda.synchronize()
    assert dts != ref_dts or not torch.equal(m.qq.data, ref_q)


@pytest.mark.gpu
@pytest.mark.parametrize("case", list(CASES))
def test_update_heated_follows_the_reference_order(ctx, case):
    """update_heated with the functor equals update, the heating, a fill and
    the CFL of the heated state in sequence (the functor runs after the
    update's last sweep and again after the heating)."""
    import torch

    from miso_amd import hip
    out = []
    for fused in (True, False):
        m, bc, _ = _setup(ctx, CASES[case])
        integ, qq = m.integrator, m.qq
        gen = torch.Generator(device="cpu").manual_seed(3)
        qrad = (0.05 * torch.randn(qq.data.shape[1:], generator=gen, dtype=torch.float64)).to(
            "cuda:0")
        dt = integ.cfl(qq)
        if fused:
            nxt = integ.update_heated(dt, bc, qq, qrad)
        else:
            integ.update(dt, bc, qq)
            hip.call("misohip_rt_heat", ctx.handle, C.byref(m.grid.c_struct()),
                     C.byref(qq.c_struct()), C.c_void_p(qrad.data_ptr()), float(dt))
            integ.apply_boundary_condition(bc, qq)
            nxt = integ.cfl(qq)
        torch.cuda.synchronize()
        out.append((nxt, qq.data.clone()))
    assert out[0][0] == out[1][0] and out[0][0] > 0
    assert torch.equal(out[0][1], out[1][1])

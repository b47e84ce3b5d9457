"""nd_rk4_noalloc, the RK4 entry point that allocates nothing, called through the C ABI:
two calls in a row (the second on a smaller network) equal nd_rk4's stage-launch path bit
for bit, and a network larger than its scratch is refused."""
import ctypes as C

import numpy as np
import pytest


def _noalloc_rk4(ND, nw, u0, p, dev, t0, dt, nsteps):
    import torch
    u = torch.from_numpy(np.array(u0)).to(dev)
    pd = torch.from_numpy(np.ascontiguousarray(p)).to(dev)
    stream = torch.cuda.current_stream(dev).cuda_stream
    ND._abi.check(ND._abi.lib().nd_rk4_noalloc(nw.plan, C.c_void_p(u.data_ptr()), C.c_void_p(pd.data_ptr()),
                                                float(t0), float(dt), int(nsteps), nw.dim, nw.pdim, stream))
    torch.cuda.synchronize()
    return u.cpu().numpy()


def _stage_rk4(nw, u0, p, dev, t0, dt, nsteps):
    import torch
    nw.tune(rk4_path=2)
    u = torch.from_numpy(np.array(u0)).to(dev)
    nw.rk4(u, torch.from_numpy(np.ascontiguousarray(p)).to(dev), t0, dt, nsteps)
    torch.cuda.synchronize()
    return u.cpu().numpy()


@pytest.mark.gpu
def test_rk4_noalloc_matches_nd_rk4(ND, dev):
    # first call: a 300k-vertex swing network (compact outputs ping-ponged through the
    # scratch, parameter copies, hubs)
    nw_a, p_a, u_a = ND.workloads.swing_ba(300_000, 2, seed=71)
    assert nw_a.tuning()["gather_mode"] == 3           # the compact-output path
    got_a = _noalloc_rk4(ND, nw_a, u_a, p_a, dev, 0.0, 0.005, 20)
    ref_a = _stage_rk4(nw_a, u_a, p_a, dev, 0.0, 0.005, 20)
    assert np.all(np.isfinite(got_a)) and np.array_equal(got_a, ref_a)
    # second call, smaller: first-order Kuramoto on G(3000, 12000), time-independent models
    nw_b, p_b, u_b = ND.workloads.kuramoto_er(3000, 12000, 3.0, seed=72)
    got_b = _noalloc_rk4(ND, nw_b, u_b, p_b, dev, 0.5, 0.01, 50)
    ref_b = _stage_rk4(nw_b, u_b, p_b, dev, 0.5, 0.01, 50)
    assert np.all(np.isfinite(got_b)) and np.array_equal(got_b, ref_b)
    # and the first network again after the smaller one
    again = _noalloc_rk4(ND, nw_a, u_a, p_a, dev, 0.0, 0.005, 20)
    assert np.array_equal(again, ref_a)


@pytest.mark.gpu
def test_rk4_noalloc_refuses_networks_beyond_its_scratch(ND, dev):
    import torch
    # 1.4e6 swing vertices: 3 x 2.8e6 stage doubles alone exceed the 2^23 of the scratch
    nw, p, u0 = ND.workloads.swing_ba(1_400_000, 2, seed=73)
    u = torch.from_numpy(u0).to(dev)
    pd = torch.from_numpy(p).to(dev)
    rc = ND._abi.lib().nd_rk4_noalloc(nw.plan, C.c_void_p(u.data_ptr()), C.c_void_p(pd.data_ptr()), 0.0, 0.01, 1,
                                      nw.dim, nw.pdim, None)
    assert rc == ND._abi.ND_ERR_ARGUMENT
    assert "scratch" in ND._abi.lib().nd_last_error().decode()
    assert torch.equal(u, torch.from_numpy(u0).to(dev))

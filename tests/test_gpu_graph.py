"""Device random stream inside the step, a whole block as one HIP graph, and
the driver: graph replays against the eager block loop, bit for bit."""
import dataclasses

import numpy as np
import pytest

from ad_afqmc_prototype_amd import synthetic as syn

pytestmark = pytest.mark.gpu


def _setup(gpu_device, nw=64, steps=5, kind="restricted", seed=0, n=24, o=6, g=96, precision="fp64"):
    from ad_afqmc_prototype_amd.core.system import system
    from ad_afqmc_prototype_amd.ham.chol import ham_chol
    from ad_afqmc_prototype_amd.meas.rhf import make_rhf_meas_ops
    from ad_afqmc_prototype_amd.prop.afqmc import init_prop_state
    from ad_afqmc_prototype_amd.prop.chol_afqmc_ops import make_chol_afqmc_ops
    from ad_afqmc_prototype_amd.prop.types import afqmc_params
    from ad_afqmc_prototype_amd.trial.rhf import make_rhf_trial_ops, rhf_trial
    s = syn.make_system(seed, n, o, g, chol_scale=0.25)
    sys_ = system(norb=n, nelec=(o, o), walker_kind=kind)
    ham = ham_chol.from_arrays(s.h0, s.h1, s.chol, device=gpu_device)
    trial = rhf_trial(s.mo_up)
    t_ops = make_rhf_trial_ops(sys_)
    m_ops = make_rhf_meas_ops(sys_, precision=precision)
    p_ops = make_chol_afqmc_ops(ham, kind, precision=precision)
    params = afqmc_params(dt=0.01, n_prop_steps=steps, n_walkers=nw, n_ortho=steps, seed=seed,
                          n_eql_blocks=5, n_blocks=10)
    kw = dict(sys=sys_, params=params, ham_data=ham, trial_data=trial, meas_ops=m_ops, prop_ops=p_ops,
              prop_ctx=p_ops.build_prop_ctx(t_ops.get_rdm1(trial), params.dt),
              meas_ctx=m_ops.build_meas_ctx(ham, trial))
    st = init_prop_state(sys=sys_, n_walkers=nw, seed=seed, ham_data=ham, trial_ops_=t_ops, trial_data=trial,
                         meas_ops=m_ops)
    return st, kw, t_ops


def test_step_draws_fields_from_the_device_key(gpu_device):
    """afqmc_step without fields == afqmc_step with normal(split(key)[1]) injected."""
    import torch
    from ad_afqmc_prototype_amd import kernels
    from ad_afqmc_prototype_amd.prop.afqmc import afqmc_step
    st, kw, _ = _setup(gpu_device)
    kw = {k: v for k, v in kw.items() if k != "sys"}
    a = afqmc_step(st, **kw)
    nk, sk = kernels.rng_split(st.rng_key)
    b = afqmc_step(st, fields=kernels.rng_normal(sk, (64, 96)), **kw)
    assert torch.equal(a.walkers, b.walkers) and torch.equal(a.weights, b.weights)
    assert torch.equal(a.rng_key, nk) and torch.equal(b.rng_key, nk)


@pytest.mark.parametrize("precision", ["fp64", "bf16x3"])
def test_captured_block_replays_match_eager(gpu_device, precision):
    """bf16x3: the state also carries the Green's function's bf16 planes (made
    by the Green's function kernel, gathered by the comb), which the graph
    copies like every other state buffer."""
    import torch
    from ad_afqmc_prototype_amd.prop.blocks import _clone_state, afqmc_block, captured_block
    st, kw, _ = _setup(gpu_device, precision=precision)
    if precision == "bf16x3":
        assert st.greens.split is not None
    ref, st_e = [], _clone_state(st)
    for _ in range(3):
        st_e, obs = afqmc_block(st_e, **kw)
        ref.append((obs.scalars["energy"].item(), obs.scalars["weight"].item(), st_e.walkers.clone(),
                    st_e.weights.clone(), st_e.rng_key.clone(), st_e.pop_control_ene_shift.item()))
    cb = captured_block(st, **kw)
    for i in range(3):
        e, w = cb.replay()
        torch.cuda.synchronize()
        e_ref, w_ref, walk_ref, wt_ref, key_ref, pop_ref = ref[i]
        assert e.item() == e_ref and w.item() == w_ref, (i, e.item(), e_ref)
        assert torch.equal(cb.state.walkers, walk_ref) and torch.equal(cb.state.weights, wt_ref)
        assert torch.equal(cb.state.rng_key, key_ref)
        assert cb.state.pop_control_ene_shift.item() == pop_ref
        # eager work of the program's own between replays
        x = torch.randn(1000, device=gpu_device).sum().item()
        assert np.isfinite(x)


def test_captured_block_from_a_stepped_bf16x3_state(gpu_device):
    """The bf16x3 step leaves only the force bias's bf16 planes in the Green's
    cache; a block captured from such a state (the cache refilled at capture)
    replays like the eager blocks from the same state."""
    import torch
    from ad_afqmc_prototype_amd.prop.afqmc import afqmc_step
    from ad_afqmc_prototype_amd.prop.blocks import _clone_state, afqmc_block, captured_block
    st, kw, _ = _setup(gpu_device, precision="bf16x3")
    st = afqmc_step(st, **{k: v for k, v in kw.items() if k != "sys"})
    assert st.greens.greens is None and st.greens.split is not None
    ref, st_e = [], _clone_state(st)
    for _ in range(2):
        st_e, obs = afqmc_block(st_e, **kw)
        ref.append((obs.scalars["energy"].item(), st_e.walkers.clone()))
    cb = captured_block(st, **kw)
    for i in range(2):
        e, _ = cb.replay()
        torch.cuda.synchronize()
        assert e.item() == ref[i][0], (i, e.item(), ref[i][0])
        assert torch.equal(cb.state.walkers, ref[i][1])


def test_driver_graph_equals_eager(gpu_device):
    from ad_afqmc_prototype_amd.driver import run_afqmc_energy
    out = []
    for use_graph in (False, True):
        st, kw, t_ops = _setup(gpu_device, steps=4)
        mu, se, e_b, w_b = run_afqmc_energy(sys=kw["sys"], params=kw["params"], ham_data=kw["ham_data"],
                                            trial_data=kw["trial_data"], meas_ops=kw["meas_ops"],
                                            trial_ops=t_ops, prop_ops=kw["prop_ops"], verbose=False,
                                            use_graph=use_graph)
        out.append((mu, se, e_b, w_b))
    (mu0, se0, e0, w0), (mu1, se1, e1, w1) = out
    assert e0.shape == (1 + 5 + 10,) and np.all(np.isfinite(e0))
    np.testing.assert_array_equal(e0, e1)
    np.testing.assert_array_equal(w0, w1)
    assert mu0 == mu1 and se0 == se1 and se0 > 0


def test_driver_rejects_too_few_blocks(gpu_device):
    from ad_afqmc_prototype_amd.driver import run_afqmc_energy
    st, kw, t_ops = _setup(gpu_device)
    with pytest.raises(ValueError):
        run_afqmc_energy(sys=kw["sys"], params=dataclasses.replace(kw["params"], n_blocks=5),
                         ham_data=kw["ham_data"], trial_data=kw["trial_data"], meas_ops=kw["meas_ops"],
                         trial_ops=t_ops, verbose=False)

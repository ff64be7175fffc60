"""The driver (driver.py:39-176 of the reference) on the HIP block against the
CPU oracle: equilibration and sampling blocks on the device Threefry stream
of the seed, against the oracle's blocks fed the same keys
(oracle/threefry.py: split per step, normals for the fields, a uniform per
comb), then outlier rejection and the blocking analysis."""
import dataclasses

import numpy as np
import pytest

from ad_afqmc_prototype_amd import synthetic as syn
from ad_afqmc_prototype_amd.stats import blocking_analysis_ratio, reject_outliers
from oracle import afqmc_oracle as orc
from oracle import threefry as tf

pytestmark = pytest.mark.gpu

N, O, G, NW, DT, STEPS, SEED = 24, 6, 96, 128, 0.01, 8, 0


def _gpu_setup(device):
    from ad_afqmc_prototype_amd.core.system import system
    from ad_afqmc_prototype_amd.ham.chol import ham_chol
    from ad_afqmc_prototype_amd.meas.rhf import make_rhf_meas_ops
    from ad_afqmc_prototype_amd.prop.chol_afqmc_ops import make_chol_afqmc_ops
    from ad_afqmc_prototype_amd.prop.types import afqmc_params
    from ad_afqmc_prototype_amd.trial.rhf import make_rhf_trial_ops, rhf_trial
    s = syn.make_system(SEED, N, O, G, chol_scale=0.25)
    sys_ = system(norb=N, nelec=(O, O), walker_kind="restricted")
    ham = ham_chol.from_arrays(s.h0, s.h1, s.chol, device=device)
    trial = rhf_trial(s.mo_up)
    params = afqmc_params(dt=DT, n_prop_steps=STEPS, n_walkers=NW, seed=SEED, n_eql_blocks=5,
                          n_blocks=10)
    kw = dict(sys=sys_, params=params, ham_data=ham, trial_data=trial,
              meas_ops=make_rhf_meas_ops(sys_), trial_ops=make_rhf_trial_ops(sys_),
              prop_ops=make_chol_afqmc_ops(ham, "restricted"))
    return s, kw


def _oracle_blocks(s, n_blocks):
    """The oracle's blocks with the fields and comb uniforms of the seed's
    stream in the reference's order (prop/afqmc.py:86-88, prop/blocks.py:81-82)."""
    sd = orc.make_system_data(s.h0, s.h1, s.chol, (s.mo_up, s.mo_dn), "restricted", DT)
    prm = orc.params(dt=DT, n_prop_steps=STEPS)
    st = orc.init_state(sd, NW, overlap_mode="log")
    key = tf.prng_key(SEED)
    es, ws, idx = [st.e_estimate], [float(NW)], []
    for _ in range(n_blocks):
        fields = []
        for _t in range(STEPS):
            key, sub = tf.split(key)
            fields.append(tf.normal(sub, (NW, G)))
        key, sub = tf.split(key)
        st, obs = orc.afqmc_block(st, fields, float(tf.uniform(sub)), sd, prm, overlap_mode="log")
        es.append(obs["energy"])
        ws.append(obs["weight"])
        idx.append(obs["sr_indices"])
    return np.array(es), np.array(ws), idx


@pytest.mark.parametrize("use_graph", [False, True])
def test_driver_matches_oracle(gpu_device, use_graph):
    from ad_afqmc_prototype_amd.driver import run_afqmc_energy
    s, kw = _gpu_setup(gpu_device)
    mu, se, e_b, w_b = run_afqmc_energy(**kw, verbose=False, use_graph=use_graph)
    e_ref, w_ref, _ = _oracle_blocks(s, 15)
    # the walk moves: no block is the clipped or fallback estimate
    assert np.all(np.abs(e_ref[1:] - e_ref[0]) > 1e-6) and np.ptp(e_ref) > 1e-3
    assert e_b.shape == e_ref.shape  # no outlier in this run
    np.testing.assert_allclose(e_b, e_ref, rtol=1e-10)
    np.testing.assert_allclose(w_b, w_ref, rtol=1e-9)
    clean, _ = reject_outliers(np.column_stack((e_ref[6:], w_ref[6:])), obs=0)
    st = blocking_analysis_ratio(clean[:, 0], clean[:, 1], print_q=False)
    assert mu == pytest.approx(st["mu"], rel=1e-10)
    assert se == pytest.approx(st["se_star"], rel=1e-6)


def test_driver_returns_outlier_cleaned_blocks(gpu_device):
    """driver.py:161-176: the sampling blocks returned are those left after
    reject_outliers; a planted outlier block is dropped from the arrays."""
    import torch
    from ad_afqmc_prototype_amd.driver import run_afqmc_energy
    from ad_afqmc_prototype_amd.prop.blocks import afqmc_block, block_obs
    s, kw = _gpu_setup(gpu_device)
    calls = {"n": 0}

    def planted(state, **bkw):
        state, obs = afqmc_block(state, **bkw)
        calls["n"] += 1
        if calls["n"] == 9:  # the fourth sampling block
            e = obs.scalars["energy"] + 1.0e3
            obs = block_obs(scalars={"energy": e, "weight": obs.scalars["weight"]})
        return state, obs

    mu, se, e_b, w_b = run_afqmc_energy(**kw, block_fn=planted, verbose=False)
    e_ref, w_ref, _ = _oracle_blocks(s, 15)
    keep = np.ones(15 + 1, dtype=bool)
    keep[9] = False
    assert e_b.shape == (15,)
    np.testing.assert_allclose(e_b, e_ref[keep], rtol=1e-10)
    np.testing.assert_allclose(w_b, w_ref[keep], rtol=1e-9)
    st = blocking_analysis_ratio(e_ref[keep][6:], w_ref[keep][6:], print_q=False)
    assert mu == pytest.approx(st["mu"], rel=1e-10)
    assert torch.cuda.is_available()

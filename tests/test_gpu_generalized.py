"""Generalized walkers (nw, 2n, 2o) in the restricted basis with an RHF trial on
the HIP path against the CPU oracle (fp64, 1e-10 relative): overlap, force
bias, Trotter step, a chain of afqmc_step calls with injected fields, QR and
comb reconfiguration (bit-exact indices).  The reference has no energy kernel
for these walkers (meas/rhf.py:169-174), so the state starts from a given
energy estimate."""
import numpy as np
import pytest

from ad_afqmc_prototype_amd import synthetic as syn
from oracle import afqmc_oracle as orc

pytestmark = pytest.mark.gpu

N, O, G, NW, DT = 14, 3, 40, 32, 0.01


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300))


def _setup(device):
    import torch
    from ad_afqmc_prototype_amd.core.system import system
    from ad_afqmc_prototype_amd.ham.chol import ham_chol
    from ad_afqmc_prototype_amd.meas.rhf import make_rhf_meas_ops
    from ad_afqmc_prototype_amd.prop.chol_afqmc_ops import make_chol_afqmc_ops
    from ad_afqmc_prototype_amd.trial.rhf import make_rhf_trial_ops, rhf_trial
    s = syn.make_system(21, N, O, G, chol_scale=0.4)
    sys_ = system(N, (O, O), "generalized")
    ham = ham_chol.from_arrays(s.h0, s.h1, s.chol, device=device)
    trial = rhf_trial(s.mo_up)
    t_ops, m_ops = make_rhf_trial_ops(sys_), make_rhf_meas_ops(sys_)
    p_ops = make_chol_afqmc_ops(ham, "generalized")
    sd = orc.make_system_data(s.h0, s.h1, s.chol, (s.mo_up, s.mo_up), "generalized", DT)
    rng = np.random.default_rng(4)
    w0 = orc.init_walkers(orc.get_rdm1(sd.trial), O, NW, "generalized")
    w0 = w0 + 0.05 * (rng.standard_normal(w0.shape) + 1j * rng.standard_normal(w0.shape))
    return dict(s=s, sys=sys_, ham=ham, trial=trial, t_ops=t_ops, m_ops=m_ops, p_ops=p_ops, sd=sd,
                w0=w0, wdev=torch.tensor(w0, device=device))


def test_overlap_force_bias_trotter(gpu_device):
    import torch
    from ad_afqmc_prototype_amd.core.ops import k_force_bias
    c = _setup(gpu_device)
    lo = c["m_ops"].log_overlap(c["wdev"], c["trial"]).cpu().numpy()
    lo_ref = orc.overlap(c["w0"], c["sd"].trial, "generalized", "log")
    assert _rel(np.exp(lo), np.exp(lo_ref)) < 1e-10
    m_ctx = c["m_ops"].build_meas_ctx(c["ham"], c["trial"])
    fb = c["m_ops"].require_kernel(k_force_bias)(c["wdev"], c["ham"], m_ctx, c["trial"]).cpu().numpy()
    fb_ref = orc.force_bias(c["w0"], c["sd"].trial, c["sd"].mctx, "generalized")
    assert _rel(fb, fb_ref) < 1e-10
    p_ctx = c["p_ops"].build_prop_ctx(c["t_ops"].get_rdm1(c["trial"]), DT)
    x = np.random.default_rng(5).standard_normal((NW, G))
    w1 = c["p_ops"].apply_trotter(c["wdev"], torch.tensor(x, device=gpu_device), p_ctx, 6).cpu().numpy()
    w1_ref = orc.apply_trotter(c["w0"], x, c["sd"].pctx, c["s"].chol, 6, "generalized")
    assert _rel(w1, w1_ref) < 1e-10
    with pytest.raises(KeyError):
        c["m_ops"].require_kernel("energy")


def test_step_chain_qr_and_comb(gpu_device):
    import torch
    from ad_afqmc_prototype_amd.prop.afqmc import afqmc_step, init_prop_state
    from ad_afqmc_prototype_amd.prop.blocks import orthonormalize_state, reconfigure
    from ad_afqmc_prototype_amd.prop.types import afqmc_params
    c = _setup(gpu_device)
    e_est = -3.0
    st = init_prop_state(sys=c["sys"], n_walkers=NW, seed=1, ham_data=c["ham"], trial_ops_=c["t_ops"],
                         trial_data=c["trial"], meas_ops=c["m_ops"], initial_walkers=c["wdev"],
                         initial_e_estimate=e_est)
    ost = orc.init_state(c["sd"], NW, overlap_mode="log", walkers=c["w0"], e_estimate=e_est)
    params = afqmc_params(dt=DT, n_walkers=NW)
    p_ctx = c["p_ops"].build_prop_ctx(c["t_ops"].get_rdm1(c["trial"]), DT)
    m_ctx = c["m_ops"].build_meas_ctx(c["ham"], c["trial"])
    prm = orc.params(dt=DT)
    for t in range(6):
        x = syn.fields_for_step(21, t, NW, G)
        st = afqmc_step(st, params=params, ham_data=c["ham"], trial_data=c["trial"], meas_ops=c["m_ops"],
                        prop_ops=c["p_ops"], prop_ctx=p_ctx, meas_ctx=m_ctx,
                        fields=torch.tensor(x, device=gpu_device))
        ost = orc.afqmc_step(ost, x, c["sd"], prm, "log")
        if t == 2:
            st = orthonormalize_state(st, sys=c["sys"], trial_data=c["trial"], meas_ops_=c["m_ops"])
            w = orc.orthonormalize(ost.walkers, "generalized")
            from dataclasses import replace
            ost = replace(ost, walkers=w, overlaps=orc.overlap(w, c["sd"].trial, "generalized", "log"))
    assert _rel(st.walkers.cpu().numpy(), ost.walkers) < 1e-10
    assert _rel(st.weights.cpu().numpy(), ost.weights) < 1e-10
    assert _rel(np.exp(st.log_overlaps.cpu().numpy() - ost.overlaps), 1.0) < 1e-10
    assert float(st.weights.min()) > 0.0
    zeta = 0.37
    st2, idx = reconfigure(st, sys=c["sys"], trial_data=c["trial"], meas_ops_=c["m_ops"], zeta=zeta)
    w_sr, w_new, idx_ref = orc.stochastic_reconfiguration(ost.walkers, st.weights.cpu().numpy(), zeta,
                                                          "generalized")
    np.testing.assert_array_equal(idx.cpu().numpy(), idx_ref)
    assert _rel(st2.walkers.cpu().numpy(), w_sr) < 1e-10


@pytest.mark.parametrize("precision,tol", [("fp32", 2e-5), ("bf16x3", 5e-4)])
def test_mixed_precision_trotter(gpu_device, precision, tol):
    import torch
    from ad_afqmc_prototype_amd.prop.chol_afqmc_ops import make_chol_afqmc_ops
    c = _setup(gpu_device)
    p_ops = make_chol_afqmc_ops(c["ham"], "generalized", precision=precision)
    p_ctx = p_ops.build_prop_ctx(c["t_ops"].get_rdm1(c["trial"]), DT)
    x = np.random.default_rng(6).standard_normal((NW, G))
    w1 = p_ops.apply_trotter(c["wdev"], torch.tensor(x, device=gpu_device), p_ctx, 6).cpu().numpy()
    w1_ref = orc.apply_trotter(c["w0"], x, c["sd"].pctx, c["s"].chol, 6, "generalized")
    assert _rel(w1, w1_ref) < tol


def test_generalized_basis_trotter(gpu_device):
    """A generalized-basis Hamiltonian (spin-orbital h1, L of size nso = 2n):
    prop ctx (no v0 / v1 correction) and the Trotter step of generalized
    walkers against the oracle (1e-10), and the reference's factory checks."""
    import torch
    from ad_afqmc_prototype_amd.ham.chol import ham_chol
    from ad_afqmc_prototype_amd.prop.chol_afqmc_ops import make_chol_afqmc_ops
    rng = np.random.default_rng(8)
    nso, g, o2, nw, dt = 2 * N, 30, 2 * O, 12, 0.01
    a = rng.standard_normal((nso, nso))
    h1 = 0.3 * (a + a.T)
    b = rng.standard_normal((g, nso, nso))
    chol = 0.1 * (b + b.transpose(0, 2, 1))
    ham = ham_chol.from_arrays(0.4, h1, chol, basis="generalized", device=gpu_device)
    with pytest.raises(ValueError):
        make_chol_afqmc_ops(ham, "restricted")
    c = np.linalg.qr(rng.standard_normal((nso, o2)))[0]
    dm = c @ c.T
    w = c[None] + 0.2 * (rng.standard_normal((nw, nso, o2)) + 1j * rng.standard_normal((nw, nso, o2)))
    x = rng.standard_normal((nw, g))
    ref_ctx = orc.build_prop_ctx(0.4, h1, chol, dm, dt, basis="generalized")
    ref = orc.apply_trotter(w, x, ref_ctx, chol, 6, "restricted")
    for prec, tol in (("fp64", 1e-10), ("fp32", 2e-5), ("bf16x3", 5e-4)):
        ops = make_chol_afqmc_ops(ham, "generalized", precision=prec)
        ctx = ops.build_prop_ctx(torch.tensor(dm, device=gpu_device), dt)
        assert _rel(ctx.mf_shifts.cpu().numpy(), ref_ctx.mf_shifts) < 1e-12
        assert abs(ctx.h0_prop - ref_ctx.h0_prop) < 1e-12 * abs(ref_ctx.h0_prop)
        out = ops.apply_trotter(torch.tensor(w, device=gpu_device), torch.tensor(x, device=gpu_device), ctx, 6)
        assert _rel(out.cpu().numpy(), ref) < tol, prec

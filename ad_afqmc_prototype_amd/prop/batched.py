"""Many independent systems advancing together: uneven orbital, occupied,
Cholesky and walker counts, each with its own Hamiltonian, RHF trial,
restricted walkers and random stream.

Every system is an ordinary run of the reference's step (prop/afqmc.py:74-147)
and block (prop/blocks.py:22-95); the batch changes how the work is launched,
not what is computed.  Each stage of the step is one grouped HIP launch over
all systems (random fields, force-bias GEMM, field shifts, VHS GEMM, Trotter,
Green's function / log-overlap, phaseless update); the fp64 -> fp32 operand
conversions run once over packed buffers.  A system's walkers, weights and
block energies equal those of running it alone with afqmc_step in fp32
precision (the same kernel bodies, the same arithmetic).

Layout: every per-system array lives at an aligned offset of one packed
buffer per quantity; two descriptor tables (device memory, written once) name
them, the second with input and output buffers swapped, so consecutive steps
alternate tables and nothing is rewritten between steps.  A block (steps, QR,
measurement, comb) can be captured once as a HIP graph and replayed.

GEMM tiles: the grouped fp32 GEMM has four tile shapes (128x128, 64x128,
128x64, 64x64 rows x columns of C per workgroup), the grouped bf16x3 GEMM
three (128x128 staged through registers, 128x256 and 256x256 moved to LDS by
buffer_load ... lds); tune() times each on the batch's own problem set with HIP
events and keeps the fastest for the force-bias and the VHS contraction
separately (optionally cached in a JSON file keyed by the precision and the
problem shapes).

In bf16x3 precision the VHS GEMM writes each system's V as split-bf16 pair
words of its upper 16 x 16 blocks (half the bytes of full fp32 rows), and the
Trotter step reads those and E - I in the same layout, as the single-system
bf16x3 step does.
"""
from __future__ import annotations

import ctypes
import json
import os
from dataclasses import dataclass
from typing import Any, Sequence

import torch

from .. import _lib
from .. import kernels
from ..core.system import system
from ..ham.chol import ham_chol
from ..meas.rhf import deferred_view, exchange_lower, make_rhf_meas_ops, quadratic_exchange
from ..trial.rhf import make_rhf_trial_ops, rhf_trial
from .afqmc import init_prop_state
from .chol_afqmc_ops import chol_sym_rows, make_chol_afqmc_ops

c128 = torch.complex128
f64 = torch.float64
f32 = torch.float32

_vp = ctypes.c_void_p


class _system_step(ctypes.Structure):
    """Host mirror of afq_system_step (include/afqmc_hip.h)."""

    _fields_ = [("n", ctypes.c_int), ("o", ctypes.c_int), ("naux", ctypes.c_int), ("nw", ctypes.c_int),
                ("ldv", ctypes.c_int), ("walker0", ctypes.c_int), ("tblock0", ctypes.c_int),
                ("ldr", ctypes.c_int),
                ("sqrt_dt", ctypes.c_double), ("dt", ctypes.c_double), ("e_estimate", ctypes.c_double),
                ("h0_prop", ctypes.c_double), ("weight_floor", ctypes.c_double),
                ("weight_cap", ctypes.c_double), ("damping", ctypes.c_double),
                ("h0", ctypes.c_double)] + [(name, _vp) for name in (
                    "trial", "walkers_in", "walkers_out", "greens", "logdet", "fields", "fb", "mf",
                    "xbar", "shift_term", "fb_term", "exp_h1_half", "vhs", "log_overlaps_in",
                    "log_overlaps_out", "weights_in", "weights_out", "pop_in", "pop_out", "key_in",
                    "key_out", "rot_chol", "rot_h1", "energy", "greens_hi", "greens_lo")] + [
                    ("ldk", ctypes.c_longlong), ("xbar_hi", _vp), ("xbar_lo", _vp), ("ldx", ctypes.c_longlong),
                ("exch_z", _vp), ("exp_h1", _vp)]


class _gemm_problem(ctypes.Structure):
    """Host mirror of afq_gemm_problem (include/afqmc_hip.h)."""

    _fields_ = [("M", ctypes.c_int), ("N", ctypes.c_int), ("K", ctypes.c_int), ("tile_start", ctypes.c_int),
                ("lda", ctypes.c_longlong), ("ldb", ctypes.c_longlong), ("ldc", ctypes.c_longlong),
                ("A", _vp), ("B", _vp), ("C", _vp), ("alpha", ctypes.c_float), ("c_fp64", ctypes.c_int),
                ("sym_n", ctypes.c_int), ("sym_ldv", ctypes.c_int), ("b_lower", ctypes.c_int)]


class _gemm_problem_bf16x3(ctypes.Structure):
    """Host mirror of afq_gemm_problem_bf16x3 (include/afqmc_hip.h)."""

    _fields_ = [("M", ctypes.c_int), ("N", ctypes.c_int), ("K", ctypes.c_int), ("tile_start", ctypes.c_int),
                ("lda", ctypes.c_longlong), ("ldb", ctypes.c_longlong), ("ldc", ctypes.c_longlong),
                ("A_hi", _vp), ("A_lo", _vp), ("B_hi", _vp), ("B_lo", _vp), ("C", _vp),
                ("alpha", ctypes.c_float), ("c_fp64", ctypes.c_int), ("sym_n", ctypes.c_int),
                ("sym_ldv", ctypes.c_int), ("b_lower", ctypes.c_int)]


class _split_problem(ctypes.Structure):
    """Host mirror of afq_split_problem (include/afqmc_hip.h)."""

    _fields_ = [("src", _vp), ("hi", _vp), ("lo", _vp), ("rows", ctypes.c_int), ("cols", ctypes.c_int),
                ("ld_src", ctypes.c_int), ("ld_dst", ctypes.c_int)]


assert ctypes.sizeof(_system_step) == 8 * 4 + 8 * 8 + 26 * 8 + 8 + 2 * 8 + 8 + 8 + 8
assert ctypes.sizeof(_gemm_problem) == 88
assert ctypes.sizeof(_gemm_problem_bf16x3) == 104
assert ctypes.sizeof(_split_problem) == 40
GEMM_SHAPES = (0, 1, 2, 3)        # grouped fp32 GEMM tile shapes
BF16X3_SHAPES = (0, 1, 2)         # grouped bf16x3 GEMM tile shapes


def gemm_tile_shape(shape: int, precision: str = "fp32") -> tuple[int, int]:
    """(rows, columns) of C per workgroup of a grouped GEMM tile shape."""
    bm, bn = ctypes.c_int(), ctypes.c_int()
    entry = "afq_gemm_bf16x3_tile_shape" if precision == "bf16x3" else "afq_gemm_tile_shape"
    _lib.call(entry, shape, ctypes.byref(bm), ctypes.byref(bn))
    return bm.value, bn.value


@dataclass(frozen=True)
class batch_system:
    """One member of the batch: Hamiltonian (restricted basis, on the device),
    RHF trial, walker count and seed of its random stream."""

    ham: ham_chol
    trial: rhf_trial
    n_walkers: int
    seed: int


class _packer:
    """Aligned sub-ranges of one flat buffer per quantity."""

    ALIGN = 64  # elements

    def __init__(self, sizes: Sequence[int], dtype, device):
        offs, o = [], 0
        for s in sizes:
            offs.append(o)
            o += (int(s) + self.ALIGN - 1) // self.ALIGN * self.ALIGN
        self.buf = torch.zeros(max(o, 1), dtype=dtype, device=device)
        self.offs = offs
        self.sizes = [int(s) for s in sizes]

    def part(self, i: int, shape=None):
        v = self.buf[self.offs[i]:self.offs[i] + self.sizes[i]]
        return v if shape is None else v.view(*shape)

    def ptr(self, i: int) -> int:
        return self.part(i).data_ptr()


def _upload(structs) -> torch.Tensor:
    raw = ctypes.string_at(ctypes.addressof(structs), ctypes.sizeof(structs))
    return torch.frombuffer(bytearray(raw), dtype=torch.uint8).clone()


class system_batch:
    """The state of every system plus the grouped step.

    Walkers, weights, log-overlaps, population shifts and keys are kept in two
    buffer sets (A, B); steps alternate A -> B and B -> A, so after an even
    number of steps the state is in A again.  QR (every ``qr_every`` steps),
    measurement and comb (every ``sr_every`` steps, prop/blocks.py:53-92) run
    on the A set; both cadences must be even.

    ``precision`` is the arithmetic of the force-bias and VHS contractions and
    of the Trotter step: "bf16x3" (split-bf16 operands on the bf16 matrix
    cores, fp32 sums) or "fp32" (fp32 matrix cores); the GEMM tile shapes are
    ``gemm_shapes`` (force bias, VHS), picked by tune().  Green's functions, QR, weights
    and the energy (fp32 products, fp64 sums) are the same in both."""

    def __init__(self, systems: Sequence[batch_system], *, dt: float, qr_every: int = 10,
                 sr_every: int = 20, n_exp_terms: int = 6, weight_floor: float = 1.0e-3,
                 weight_cap: float = 100.0, damping: float = 0.1, shift_ema: float = 0.1,
                 gemm_shapes: tuple[int, int] | None = None, precision: str = "bf16x3",
                 merge_one_body_halves: bool = True, grouped_block_end: bool = True):
        if not systems:
            raise ValueError("empty batch")
        if precision not in ("bf16x3", "fp32"):
            raise ValueError(f"precision must be 'bf16x3' or 'fp32', not {precision!r}")
        self.precision = precision
        # bf16x3: a step followed by another leaves its last one-body half to it
        # (prop/afqmc.py afqmc_step(defer_final=True)); between such steps the
        # walkers are E^-1 W, their Green's functions are taken with the trial
        # E C and the force bias with C^T L E
        self.merge = bool(merge_one_body_halves) and precision == "bf16x3"
        # the block end of every system in two launches (afq_block_end_grouped,
        # afq_reconfigure_gather_grouped) instead of a dozen per system
        self.grouped_block_end = bool(grouped_block_end)
        self.deferred = False
        if qr_every % 2 or sr_every % 2 or qr_every <= 0 or sr_every <= 0:
            raise ValueError("qr_every and sr_every must be positive and even")
        dev = systems[0].ham.h1.device
        if dev.type != "cuda":
            raise ValueError("the batch runs on a HIP device only")
        self.device = dev
        self.dt, self.sqrt_dt = float(dt), float(dt) ** 0.5
        self.qr_every, self.sr_every = qr_every, sr_every
        self.n_exp_terms = n_exp_terms
        self.shift_ema = shift_ema
        self.floor, self.cap, self.damping = weight_floor, weight_cap, damping
        self.count = len(systems)
        self.members = []
        for s in systems:
            n, o, g = s.ham.h1.shape[0], s.trial.nocc, s.ham.chol.shape[0]
            if s.ham.basis != "restricted" or not s.trial.restricted:
                raise ValueError("batched systems need a restricted basis and an RHF trial")
            if s.trial.mo_coeff.is_complex():
                raise ValueError("batched systems need real trial orbitals")
            if o > 32 or n > 208 or o > n:
                raise ValueError(f"batched system n={n} o={o} outside n <= 208, o <= 32")
            if s.n_walkers * g >= 2 ** 31:
                raise ValueError("n_walkers * naux must stay below 2^31")
            self.members.append(self._member(s, n, o, g))
        self.max_n = max(m["n"] for m in self.members)
        self.max_o = max(m["o"] for m in self.members)
        self.total_walkers = sum(m["nw"] for m in self.members)
        self._allocate()
        self._tables()
        self.gemm_shapes = tuple(gemm_shapes) if gemm_shapes is not None else (0, 0)
        valid = BF16X3_SHAPES if precision == "bf16x3" else GEMM_SHAPES
        if len(self.gemm_shapes) != 2 or any(x not in valid for x in self.gemm_shapes):
            raise ValueError(f"gemm_shapes {self.gemm_shapes} not two of {valid} ({precision})")
        self.steps_done = 0
        self.block_stats = torch.zeros((self.count, 2), dtype=f64, device=dev)
        if self.precision == "bf16x3":
            self._split("fb")  # planes of the initial walkers' Green's functions

    # ------------------------------------------------------------------ setup
    def _member(self, s: batch_system, n: int, o: int, g: int) -> dict:
        sys_ = system(n, (o, o), "restricted")
        t_ops = make_rhf_trial_ops(sys_)
        m_ops = make_rhf_meas_ops(sys_, precision="fp32")
        p_ops = make_chol_afqmc_ops(s.ham, "restricted", precision="fp32")
        p_ctx = p_ops.build_prop_ctx(t_ops.get_rdm1(s.trial), self.dt)
        m_ctx = m_ops.build_meas_ctx(s.ham, s.trial)
        st = init_prop_state(sys=sys_, n_walkers=s.n_walkers, seed=s.seed, ham_data=s.ham,
                             trial_ops_=t_ops, trial_data=s.trial, meas_ops=m_ops)
        if self.precision == "bf16x3":
            # split planes of the half-rotated Cholesky in the bf16x3 energy
            # layout: (naux, round16(o), round32(n)) per plane, hi then lo
            opad, ldr = (o + 15) // 16 * 16, (n + 31) // 32 * 32
            pad = torch.zeros((g, opad, ldr), dtype=f64, device=s.ham.h1.device)
            pad[:, :o, :n] = m_ctx.rot_chol[0]
            rot_pad = torch.stack(kernels.split_bf16(pad.reshape(-1, ldr))).contiguous()
            del pad
            rot_h1 = m_ctx.rot_h1[:1].contiguous()
        else:
            ldr = kernels.ldv_for(n)
            rot_pad, rot_h1 = m_ctx.energy_operands(1)
        if self.precision == "bf16x3":
            ops = dict(chol_hl=kernels.split_bf16(chol_sym_rows(s.ham.chol)),
                       rot_hl=kernels.split_bf16(m_ctx.rot_chol_flat),
                       e_pairs=p_ctx.exp_h1_half_sym_pairs.contiguous())
            if self.merge:
                # the deferred walkers' operands: trial E C, C^T L E, E^2 - I
                trial_e, ctx_e = deferred_view(m_ctx, s.trial, p_ctx.exp_h1_half)
                ops.update(trial_e=trial_e.mo_coeff.contiguous(),
                           rot_e_hl=kernels.split_bf16(ctx_e.rot_chol_flat.contiguous()),
                           e2_pairs=p_ctx.exp_h1_sym_pairs.contiguous())
        else:
            ops = dict(chol_s32=kernels.convert_f32(chol_sym_rows(s.ham.chol)), rot32=m_ctx.fb_operand(1))
        if quadratic_exchange(n, o, g):
            # the exchange as a quadratic form: L, Y's lower triangle with its
            # diagonal halved (Z = 2 G L^T), in the batch's precision
            y = exchange_lower(m_ctx.rot_chol[0])
            if self.precision == "bf16x3":
                ops["y_hl"] = kernels.split_bf16(y)
            else:
                ops["y32"] = kernels.convert_f32(y)
            del y
        return dict(n=n, o=o, g=g, nw=s.n_walkers, ldv=kernels.ldv_for(n), nsym=kernels.sym_blocks(n) * 256,
                    h0=float(s.ham.h0),
                    trial=s.trial.mo_up.contiguous(), **ops, rot_pad=rot_pad, rot_h1=rot_h1, ldr=ldr,
                    e_half=p_ctx.exp_h1_half_pad32.contiguous(), mf=p_ctx.mf_shifts.contiguous(),
                    h0_prop=float(complex(p_ctx.h0_prop).real), e_est=float(st.e_estimate), state0=st)

    def _allocate(self):
        dev, ms = self.device, self.members
        P = lambda sizes, dt: _packer(sizes, dt, dev)  # noqa: E731
        self.walkers = [P([m["nw"] * m["n"] * m["o"] for m in ms], c128) for _ in range(2)]
        self.greens = P([2 * m["nw"] * m["o"] * m["n"] for m in ms], f64)
        pad8 = kernels.pad8
        if self.precision == "bf16x3":
            i16 = torch.int16
            self.greens_hl = [P([2 * m["nw"] * pad8(m["o"] * m["n"]) for m in ms], i16) for _ in range(2)]
            self.xbar_hl = [P([2 * m["nw"] * pad8(m["g"]) for m in ms], i16) for _ in range(2)]
        else:
            self.greens32 = P(self.greens.sizes, f32)
        self.logdet = P([m["nw"] for m in ms], c128)
        self.fields = P([m["nw"] * m["g"] for m in ms], f64)
        self.fb = P([2 * m["nw"] * m["g"] for m in ms], f64)
        # block ends with the exchange as a quadratic form (every system
        # n < naux): Z = G Y rows
        self.quadratic = all("y32" in m or "y_hl" in m for m in ms)
        if self.quadratic:
            self.exch_z = P([2 * m["nw"] * m["o"] * m["n"] for m in ms], f64)
        self.xbar = P([2 * m["nw"] * m["g"] for m in ms], f64)
        if self.precision == "fp32":
            self.xbar32 = P(self.xbar.sizes, f32)
        self.shift = P([m["nw"] for m in ms], c128)
        self.fbt = P([m["nw"] for m in ms], c128)
        if self.precision == "bf16x3":
            # V as pair words of its upper blocks; E - I likewise
            self.vhs = P([2 * m["nw"] * m["nsym"] for m in ms], torch.int32)
            self.e_half = P([m["nsym"] for m in ms], torch.int32)
            if self.merge:
                self.e_full = P([m["nsym"] for m in ms], torch.int32)
                self.trials_e = P([m["n"] * m["o"] for m in ms], f64)
        else:
            self.vhs = P([2 * m["nw"] * m["n"] * m["ldv"] for m in ms], f32)
            self.e_half = P([m["n"] * m["ldv"] for m in ms], f32)
        self.eloc = P([m["nw"] for m in ms], c128)
        self.lo = [P([m["nw"] for m in ms], c128) for _ in range(2)]
        self.weights = [P([m["nw"] for m in ms], f64) for _ in range(2)]
        self.pop = [torch.zeros((self.count,), dtype=f64, device=dev) for _ in range(2)]
        self.keys = [torch.zeros((self.count, 2), dtype=torch.int32, device=dev) for _ in range(2)]
        if self.grouped_block_end:
            self.comb_cw = torch.empty((sum(m["nw"] for m in ms),), dtype=f64, device=dev)
            self.comb_idx = torch.empty((sum(m["nw"] for m in ms),), dtype=torch.int32, device=dev)
            self.greens_tmp = torch.zeros_like(self.greens.buf)  # the gaps between systems stay zero
        self.mf = P([m["g"] for m in ms], c128)
        self.trials = P([m["n"] * m["o"] for m in ms], f64)
        for i, m in enumerate(ms):
            st = m.pop("state0")
            self.walkers[0].part(i).copy_(st.walkers.reshape(-1))
            self.greens.part(i).copy_(st.greens.greens.reshape(-1))
            self.lo[0].part(i).copy_(st.log_overlaps)
            self.weights[0].part(i).copy_(st.weights)
            self.pop[0][i] = st.pop_control_ene_shift
            self.keys[0][i] = st.rng_key
            self.e_half.part(i).copy_((m["e_pairs"] if self.precision == "bf16x3" else m["e_half"]).reshape(-1))
            self.mf.part(i).copy_(m["mf"])
            self.trials.part(i).copy_(m["trial"].reshape(-1))
            if self.merge:
                self.e_full.part(i).copy_(m["e2_pairs"].reshape(-1))
                self.trials_e.part(i).copy_(m["trial_e"].reshape(-1))

    def _desc(self, parity: int, idx: Sequence[int] | None = None, trial_e: bool = False):
        """Descriptor table of the systems ``idx`` (all by default) for steps
        from buffer set ``parity``; walker0 / tblock0 are prefix sums over the
        table.  ``trial_e``: the trial is E C (walkers without their last
        one-body half).  Returns (array, walkers, Trotter workgroups)."""
        a, b = parity, 1 - parity
        idx = list(range(self.count)) if idx is None else list(idx)
        arr = (_system_step * len(idx))()
        w0 = t0 = 0
        for d, i in zip(arr, idx):
            m = self.members[i]
            d.n, d.o, d.naux, d.nw, d.ldv, d.ldr = m["n"], m["o"], m["g"], m["nw"], m["ldv"], m["ldr"]
            d.walker0, d.tblock0 = w0, t0
            w0 += m["nw"]
            t0 += m["nw"] * ((m["o"] + 15) // 16)
            d.sqrt_dt, d.dt, d.e_estimate, d.h0_prop = self.sqrt_dt, self.dt, m["e_est"], m["h0_prop"]
            d.weight_floor, d.weight_cap, d.damping, d.h0 = self.floor, self.cap, self.damping, m["h0"]
            d.trial = (self.trials_e if trial_e else self.trials).ptr(i)
            d.walkers_in, d.walkers_out = self.walkers[a].ptr(i), self.walkers[b].ptr(i)
            d.greens, d.logdet = self.greens.ptr(i), self.logdet.ptr(i)
            d.fields, d.fb, d.mf, d.xbar = self.fields.ptr(i), self.fb.ptr(i), self.mf.ptr(i), self.xbar.ptr(i)
            d.shift_term, d.fb_term = self.shift.ptr(i), self.fbt.ptr(i)
            d.exp_h1_half, d.vhs = self.e_half.ptr(i), self.vhs.ptr(i)
            if self.merge:
                d.exp_h1 = self.e_full.ptr(i)
            d.log_overlaps_in, d.log_overlaps_out = self.lo[a].ptr(i), self.lo[b].ptr(i)
            d.weights_in, d.weights_out = self.weights[a].ptr(i), self.weights[b].ptr(i)
            d.pop_in, d.pop_out = self.pop[a][i:i + 1].data_ptr(), self.pop[b][i:i + 1].data_ptr()
            d.key_in, d.key_out = self.keys[a][i].data_ptr(), self.keys[b][i].data_ptr()
            d.rot_chol, d.rot_h1, d.energy = (m["rot_pad"].data_ptr(), m["rot_h1"].data_ptr(),
                                              self.eloc.ptr(i))
            if self.quadratic:
                d.exch_z = self.exch_z.ptr(i)
            if self.precision == "bf16x3":
                d.greens_hi, d.greens_lo = self.greens_hl[0].ptr(i), self.greens_hl[1].ptr(i)
                d.ldk = kernels.pad8(m["o"] * m["n"])
                d.xbar_hi, d.xbar_lo = self.xbar_hl[0].ptr(i), self.xbar_hl[1].ptr(i)
                d.ldx = kernels.pad8(m["g"])
        return arr, w0, t0

    def _problems(self, which: str, shape: int):
        bm, bn = gemm_tile_shape(shape)
        arr = (_gemm_problem * self.count)()
        start = 0
        for i, m in enumerate(self.members):
            p = arr[i]
            if which == "fb":
                k = m["o"] * m["n"]
                p.M, p.N, p.K = 2 * m["nw"], m["g"], k
                p.lda, p.ldb, p.ldc = k, m["rot32"].shape[1], m["g"]
                p.A, p.B, p.C = self.greens32.ptr(i), m["rot32"].data_ptr(), self.fb.ptr(i)
                p.alpha, p.c_fp64 = 2.0, 1
            elif which == "exch":
                k = m["o"] * m["n"]
                p.M, p.N, p.K = 2 * m["nw"], k, k
                p.lda, p.ldb, p.ldc = k, k, k
                p.A, p.B, p.C = self.greens32.ptr(i), m["y32"].data_ptr(), self.exch_z.ptr(i)
                p.alpha, p.c_fp64, p.b_lower = 2.0, 1, 1  # Z = 2 G L^T, L lower triangular
            else:
                # symmetric VHS: the upper 16 x 16 blocks, each entry stored twice
                # into the (n, ldv) rows the Trotter reads
                p.M, p.N, p.K = 2 * m["nw"], m["chol_s32"].shape[0], m["g"]
                p.lda, p.ldb, p.ldc = m["g"], m["g"], 0
                p.A, p.B, p.C = self.xbar32.ptr(i), m["chol_s32"].data_ptr(), self.vhs.ptr(i)
                p.alpha, p.c_fp64 = 1.0, 0
                p.sym_n, p.sym_ldv = m["n"], m["ldv"]
            p.tile_start = start
            start += -(-p.M // bm) * -(-p.N // bn)
        return arr, start

    def _problems_bf16x3(self, which: str, shape: int):
        pad8 = kernels.pad8
        bm, bn = gemm_tile_shape(shape, "bf16x3")
        arr = (_gemm_problem_bf16x3 * self.count)()
        start = 0
        for i, m in enumerate(self.members):
            p = arr[i]
            if which in ("fb", "fb_e"):
                k = pad8(m["o"] * m["n"])
                (bh, bl), (ah, al) = m["rot_hl" if which == "fb" else "rot_e_hl"], self.greens_hl
                p.M, p.N, p.K = 2 * m["nw"], m["g"], k
                p.ldc, p.C, p.alpha, p.c_fp64 = m["g"], self.fb.ptr(i), 2.0, 1
            elif which == "exch":
                k = pad8(m["o"] * m["n"])
                (bh, bl), (ah, al) = m["y_hl"], self.greens_hl
                p.M, p.N, p.K = 2 * m["nw"], m["o"] * m["n"], k
                p.ldc, p.C, p.alpha, p.c_fp64 = m["o"] * m["n"], self.exch_z.ptr(i), 2.0, 1
                p.b_lower = 1  # Z = 2 G L^T, L lower triangular
            else:
                # V's upper blocks as pair words, stored as they come
                k = pad8(m["g"])
                (bh, bl), (ah, al) = m["chol_hl"], self.xbar_hl
                p.M, p.N, p.K = 2 * m["nw"], bh.shape[0], k
                p.ldc, p.C, p.alpha, p.c_fp64 = m["nsym"], self.vhs.ptr(i), 1.0, 2
                assert bh.shape[0] == m["nsym"]
            assert bh.shape[1] == k and bl.shape[1] == k
            # buffer descriptors of the LDS-DMA shapes address each plane in 32 bits
            assert max(p.M, p.N) * k * 2 < 2 ** 31, "operand planes must stay below 2 GiB"
            p.lda = p.ldb = k
            p.A_hi, p.A_lo, p.B_hi, p.B_lo = ah.ptr(i), al.ptr(i), bh.data_ptr(), bl.data_ptr()
            p.tile_start = start
            start += -(-p.M // bm) * -(-p.N // bn)
        return arr, start

    def _split_table(self, which: str):
        """afq_split_bf16_grouped problems: the Green's-function rows (force
        bias) or the shifted-field rows (VHS) of every system, to padded planes."""
        arr = (_split_problem * self.count)()
        biggest = 0
        for i, m in enumerate(self.members):
            q = arr[i]
            src, (hi, lo), cols = ((self.greens, self.greens_hl, m["o"] * m["n"]) if which == "fb"
                                   else (self.xbar, self.xbar_hl, m["g"]))
            q.src, q.hi, q.lo = src.ptr(i), hi.ptr(i), lo.ptr(i)
            q.rows, q.cols, q.ld_src, q.ld_dst = 2 * m["nw"], cols, cols, kernels.pad8(cols)
            biggest = max(biggest, q.rows * q.ld_dst)
        return _upload(arr).to(self.device), biggest

    def _tables(self):
        full = [self._desc(0), self._desc(1)]
        self.desc_host = [f[0] for f in full]
        self.desc = [_upload(d).to(self.device) for d in self.desc_host]
        # Trotter launches per bucket of equal row-tile count (the workgroup is
        # sized for its bucket's largest n), energy launches per bucket of equal
        # occupied-tile count (the rotated-Cholesky layout of the system)
        def buckets(key):
            out = {}
            for i, m in enumerate(self.members):
                out.setdefault(key(m), []).append(i)
            return [out[k] for k in sorted(out)]
        # (bf16x3: one bucket per pair of row tiles, the bf16x3 kernel's wave count)
        self.trotter_buckets = []
        tiles = ((lambda m: ((m["n"] + 15) // 16 + 1) // 2) if self.precision == "bf16x3"
                 else (lambda m: (m["n"] + 15) // 16))
        for idx in buckets(tiles):
            tabs = [self._desc(p, idx) for p in (0, 1)]
            self.trotter_buckets.append(dict(
                tables=[_upload(t[0]).to(self.device) for t in tabs], count=len(idx),
                blocks=tabs[0][2], max_n=max(self.members[i]["n"] for i in idx)))
        # Green's-function launches per bucket of the kernel's occupied-orbital
        # capacity (8, 16 or 32: its per-wave LDS and registers scale with it)
        self.greens_buckets = []
        for idx in buckets(lambda m: 8 if m["o"] <= 8 else 16 if m["o"] <= 16 else 32):
            tabs = [self._desc(p, idx) for p in (0, 1)]
            self.greens_buckets.append(dict(
                tables=[_upload(t[0]).to(self.device) for t in tabs], count=len(idx), walkers=tabs[0][1],
                max_o=max(self.members[i]["o"] for i in idx)))
        # bf16x3 steps: the Gauss-Jordan kernel writing the force bias's planes
        # (capacity 16 or 32 occupied orbitals)
        self.greens_split_buckets = []
        if self.precision == "bf16x3":
            for idx in buckets(lambda m: 16 if m["o"] <= 16 else 32):
                tabs = [self._desc(p, idx) for p in (0, 1)]
                bk = dict(tables=[_upload(t[0]).to(self.device) for t in tabs], count=len(idx),
                          walkers=tabs[0][1], max_o=max(self.members[i]["o"] for i in idx))
                if self.merge:
                    bk["tables_e"] = [_upload(self._desc(p, idx, trial_e=True)[0]).to(self.device) for p in (0, 1)]
                self.greens_split_buckets.append(bk)
        self.energy_buckets = []
        for idx in buckets(lambda m: (m["o"] + 15) // 16):
            arr, nw, _ = self._desc(0, idx)
            max_n = max(self.members[i]["n"] for i in idx)
            nt = (self.members[idx[0]]["o"] + 15) // 16
            blocks, wpb = nw, 1
            if self.precision == "bf16x3":
                # two walkers per workgroup (half the rotated-Cholesky reads)
                # when their Green's-function planes fit in LDS; tblock0 then
                # counts the energy workgroups
                wpb = 2 if _lib.load().afq_energy_bf16x3_grouped_fits(max_n, nt, 2) else 1
                blocks = 0
                for d, i in zip(arr, idx):
                    d.tblock0 = blocks
                    blocks += -(-self.members[i]["nw"] // wpb)
            self.energy_buckets.append(dict(table=_upload(arr).to(self.device), count=len(idx),
                                            walkers=nw, blocks=blocks, wpb=wpb, max_n=max_n, nt=nt))
        self.gemm_tables = {}
        self.split_tables = {}
        for which in ("fb", "vhs") + (("exch",) if self.quadratic else ()) + (("fb_e",) if self.merge else ()):
            if self.precision == "bf16x3":
                for shape in BF16X3_SHAPES:
                    arr, tiles = self._problems_bf16x3(which, shape)
                    self.gemm_tables[(which, shape)] = (_upload(arr).to(self.device), tiles)
                if which == "fb":  # (the field shifts write the VHS operand's planes)
                    self.split_tables[which] = self._split_table(which)
                continue
            for shape in GEMM_SHAPES:
                arr, tiles = self._problems(which, shape)
                self.gemm_tables[(which, shape)] = (_upload(arr).to(self.device), tiles)
        if self.greens.buf.numel() >= 2 ** 31 or self.xbar.buf.numel() >= 2 ** 31:
            raise ValueError("batch too large for one conversion launch")

    # ------------------------------------------------------------------ stages
    def _split(self, which: str):
        """The bf16 planes of the force bias's (Green's-function rows) or the
        VHS's (shifted fields) A operand, from their fp64 rows."""
        stab, biggest = self.split_tables[which]
        kernels._call("afq_split_bf16_grouped", stab, self.count, stab.data_ptr(), biggest,
                      kernels._stream(stab), tag=which)

    def _grouped_gemm(self, which: str):
        if self.precision == "bf16x3":
            # the A operand's bf16 planes come from the Green's stage (force
            # bias) and the field shifts (VHS)
            shape = self.gemm_shapes[0 if which in ("fb", "exch", "fb_e") else 1]
            tab, tiles = self.gemm_tables[(which, shape)]
            kernels._call("afq_gemm_bf16x3_nt_grouped", tab, self.count, tab.data_ptr(), tiles, shape,
                          kernels._stream(tab), tag="fb" if which == "fb_e" else which)
            return
        # (the exchange's G Y has M = 2 nw like the force bias: the force bias's shape)
        shape = self.gemm_shapes[0 if which in ("fb", "exch") else 1]
        tab, tiles = self.gemm_tables[(which, shape)]
        kernels._call("afq_sgemm_nt_grouped", tab, self.count, tab.data_ptr(), tiles, shape,
                      kernels._stream(tab), tag=which)

    def _convert(self, src: _packer, dst: _packer):
        n = src.buf.numel()
        kernels._call("afq_convert_f32", src.buf, 1, n, src.buf.data_ptr(), n, dst.buf.data_ptr(), n,
                      kernels._stream(src.buf))

    def _greens(self, p: int, st: int, rows: bool = True, trial_e: bool = False):
        """Green's functions and log-overlaps of every system's new walkers
        (buffer set p), one launch per occupied-orbital bucket: fp64 rows
        (and, in bf16x3 precision, their planes split from them), or with
        rows=False (bf16x3 steps) only the force bias's planes (trial_e: of
        walkers without their last one-body half, with the trial E C)."""
        if not rows:
            for bk in self.greens_split_buckets:
                t = bk["tables_e" if trial_e else "tables"][p]
                kernels._call("afq_overlap_greens_bf16x3_grouped", t, bk["count"], t.data_ptr(), bk["walkers"],
                              bk["max_o"], st)
            return
        for bk in self.greens_buckets:
            t = bk["tables"][p]
            kernels._call("afq_overlap_greens_grouped", t, bk["count"], t.data_ptr(), bk["walkers"], bk["max_o"], st)
        if self.precision == "bf16x3":
            self._split("fb")

    def step(self, last_half: bool = True):
        """One step of every system (fields drawn from each system's key).
        last_half=False (merging batches) leaves the step's last one-body half
        to the next step; the batch is then ``deferred`` until a step runs
        with last_half=True."""
        p = self.steps_done % 2
        d = self.desc[p]
        st = kernels._stream(d)
        first_full = self.deferred
        last_half = bool(last_half) or not self.merge
        kernels._call("afq_rng_fields_grouped", d, self.count, d.data_ptr(), self.total_walkers, st)
        if self.precision == "fp32":
            self._convert(self.greens, self.greens32)
        self._grouped_gemm("fb_e" if first_full else "fb")
        kernels._call("afq_field_shifts_grouped", d, self.count, d.data_ptr(), self.total_walkers, st)
        if self.precision == "fp32":
            self._convert(self.xbar, self.xbar32)
        self._grouped_gemm("vhs")
        for bk in self.trotter_buckets:
            t = bk["tables"][p]
            if first_full or not last_half:
                kernels._call("afq_trotter_bf16x3_sym_grouped_halves", t, bk["count"], t.data_ptr(), bk["blocks"],
                              bk["max_n"], self.n_exp_terms, int(first_full), int(last_half), st,
                              tag=None if last_half else "deferred")
                continue
            kernels._call("afq_trotter_bf16x3_sym_grouped" if self.precision == "bf16x3" else "afq_trotter_f32_grouped",
                          t, bk["count"], t.data_ptr(), bk["blocks"], bk["max_n"], self.n_exp_terms, st)
        self._greens(p, st, rows=self.precision != "bf16x3", trial_e=not last_half)
        kernels._call("afq_phaseless_update_grouped", d, self.count, d.data_ptr(), st)
        self.steps_done += 1
        self.deferred = not last_half

    def _orthonormalize(self):
        """QR of every system (walkers.py:169-211) and the overlap refresh
        (prop/blocks.py:53-57); the state stays in set A."""
        if self.deferred:
            raise RuntimeError("QR of walkers without their last one-body half")
        if self.steps_done % 2:
            raise RuntimeError("QR needs the state in buffer set A (an even step count)")
        d = self.desc[0]
        st = kernels._stream(d)
        kernels._call("afq_orthonormalize_grouped", d, self.count, d.data_ptr(), self.total_walkers,
                      self.max_n, self.max_o, st)
        self._greens(0, st)
        self.walkers[0].buf.copy_(self.walkers[1].buf)
        torch.mul(self.logdet.buf, 2.0, out=self.lo[0].buf)

    def _measure_reconfigure(self):
        """Per system: clipped local energies, block energy and shift EMA
        (prop/blocks.py:59-79), comb with a uniform from the system's key
        (prop/blocks.py:81-92).  Called right after _orthonormalize.  With
        grouped_block_end every system's block end is one launch and the
        gather another; otherwise a dozen launches per system (the same
        results bit for bit)."""
        thresh = (2.0 / self.dt) ** 0.5
        if self.quadratic:
            # c (the force bias of the measured walkers), Z = G Y, then the
            # energies (afq_energy_assemble_grouped, c = fb / 2)
            d = self.desc[0]
            if self.precision == "fp32":
                self._convert(self.greens, self.greens32)
            self._grouped_gemm("fb")
            self._grouped_gemm("exch")
            kernels._call("afq_energy_assemble_grouped", d, self.count, d.data_ptr(), self.total_walkers,
                          kernels._stream(d))
        for bk in ([] if self.quadratic else self.energy_buckets):
            t = bk["table"]
            if self.precision == "bf16x3":
                kernels._call("afq_energy_bf16x3_grouped", t, bk["count"], t.data_ptr(), bk["blocks"],
                              bk["max_n"], bk["nt"], bk["wpb"], kernels._stream(t))
            else:
                kernels._call("afq_energy_f32_grouped", t, bk["count"], t.data_ptr(), bk["walkers"],
                              bk["max_n"], bk["nt"], kernels._stream(t))
        if self.grouped_block_end:
            d = self.desc[0]
            st = kernels._stream(d)
            kernels._call("afq_block_end_grouped", d, self.count, d.data_ptr(), thresh, float(self.shift_ema),
                          self.block_stats.data_ptr(), self.comb_cw.data_ptr(), self.comb_idx.data_ptr(), st)
            kernels._call("afq_reconfigure_gather_grouped", d, self.count, d.data_ptr(), self.total_walkers,
                          self.comb_idx.data_ptr(), self.greens.buf.data_ptr(), self.greens_tmp.data_ptr(), st)
            self.greens.buf.copy_(self.greens_tmp)
            if self.precision == "bf16x3":
                self._split("fb")
            return
        for i, m in enumerate(self.members):
            nw, n, o = m["nw"], m["n"], m["o"]
            g = self.greens.part(i, (2, nw, o * n))
            w = self.weights[0].part(i)
            pop = self.pop[0][i]
            e = self.eloc.part(i)
            out = kernels.block_energy(e, w, pop, e_ref=m["e_est"], thresh=thresh,
                                       shift_ema=self.shift_ema)
            self.block_stats[i].copy_(out[:2])
            pop.copy_(out[2])
            key_next, sub = kernels.rng_split(self.keys[0][i])
            self.keys[0][i].copy_(key_next)
            idx, w_new = kernels.comb_indices(w, kernels.rng_uniform(sub))
            w.copy_(w_new)
            kernels.gather_rows(self.walkers[1].part(i, (nw, n, o)), idx,
                                out=self.walkers[0].part(i, (nw, n, o)))
            g.copy_(kernels.gather_rows(g, idx, nplanes=2))
            torch.mul(kernels.gather_rows(self.logdet.part(i), idx), 2.0, out=self.lo[0].part(i))
        if self.precision == "bf16x3":
            self._split("fb")  # the planes of the surviving walkers' rows

    def advance(self, n_steps: int):
        """n_steps steps of every system with the QR / measurement cadence
        (counted from the batch's first step)."""
        for k in range(n_steps):
            t = self.steps_done + 1
            self.step(last_half=(t % self.sr_every == 0 or t % self.qr_every == 0 or k == n_steps - 1))
            t = self.steps_done
            if t % self.sr_every == 0:
                self._orthonormalize()
                self._measure_reconfigure()
            elif t % self.qr_every == 0:
                self._orthonormalize()

    # --------------------------------------------------------------- queries
    def system_state(self, i: int) -> dict:
        """Copies of system i's walkers, weights, log-overlaps, pop shift and key
        (state in set A: call after an even number of steps)."""
        if self.deferred:
            raise RuntimeError("the walkers lack their last one-body half: run a step with last_half=True")
        m = self.members[i]
        a = self.steps_done % 2
        return dict(walkers=self.walkers[a].part(i, (m["nw"], m["n"], m["o"])).clone(),
                    weights=self.weights[a].part(i).clone(), log_overlaps=self.lo[a].part(i).clone(),
                    pop=self.pop[a][i].clone(), key=self.keys[a][i].clone())

    def flops_per_step(self) -> float:
        """Algorithmic flops of one step of every system (real flops; complex x
        real MAC = 4, complex x complex MAC = 8; force bias, VHS (its n (n + 1) / 2
        distinct entries), Taylor, the two one-body halves, C^T w, LU and the
        Green's solve)."""
        tot = 0.0
        for m in self.members:
            n, o, g, nw = m["n"], m["o"], m["g"], m["nw"]
            tot += nw * (2 * g * n * (n + 1) + 4 * g * o * n + self.n_exp_terms * 8 * n * n * o
                         + 2 * 4 * n * n * o + 4 * o * o * n + 8 / 3 * o ** 3 + 8 * o * o * n)
        return tot

    # ------------------------------------------------------------- autotuner
    def signature(self) -> str:
        return self.precision + ":" + ";".join(f"{m['n']},{m['o']},{m['g']},{m['nw']}" for m in self.members)

    def _launch_gemm(self, which: str, shape: int, stream) -> None:
        tab, tiles = self.gemm_tables[(which, shape)]
        if self.precision == "bf16x3":
            _lib.call("afq_gemm_bf16x3_nt_grouped", self.count, tab.data_ptr(), tiles, shape, stream.cuda_stream)
        else:
            _lib.call("afq_sgemm_nt_grouped", self.count, tab.data_ptr(), tiles, shape, stream.cuda_stream)

    def tune(self, reps: int = 5, cache_path: str | None = None) -> dict:
        """Pick the grouped-GEMM tile shape of the force-bias and the VHS
        contraction by timing each shape of the batch's precision on this
        batch (HIP events, ``reps`` launches after one warm-up; the operands
        are whatever the buffers hold, the outputs are overwritten).  With
        ``cache_path`` the choice is read from / written to a JSON file keyed
        by the precision and the problem shapes."""
        key = self.signature()
        cache = {}
        if cache_path and os.path.exists(cache_path):
            with open(cache_path) as f:
                cache = json.load(f)
            if key in cache:
                self.gemm_shapes = tuple(cache[key]["shapes"])
                return cache[key]
        stream = torch.cuda.current_stream(self.device)
        shapes = BF16X3_SHAPES if self.precision == "bf16x3" else GEMM_SHAPES
        timings = {}
        best = []
        for which in ("fb", "vhs"):
            ms = {}
            for shape in shapes:
                self._launch_gemm(which, shape, stream)
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                for _ in range(reps):
                    self._launch_gemm(which, shape, stream)
                e1.record(stream)
                e1.synchronize()
                ms[shape] = e0.elapsed_time(e1) / reps
            timings[which] = ms
            best.append(min(ms, key=ms.get))
        self.gemm_shapes = tuple(best)
        entry = {"shapes": list(best), "precision": self.precision,
                 "ms": {w: {str(k): v for k, v in t.items()} for w, t in timings.items()}}
        if cache_path:
            cache[key] = entry
            os.makedirs(os.path.dirname(os.path.abspath(cache_path)), exist_ok=True)
            with open(cache_path, "w") as f:
                json.dump(cache, f, indent=1)
        return entry


class captured_batch_block:
    """sr_every steps of the batch (one block: steps, QR, measurement, comb)
    recorded once as a HIP graph; replay() advances every system by one block
    and leaves the block energies and weights in batch.block_stats."""

    def __init__(self, batch: system_batch, warmup_blocks: int = 1):
        if batch.steps_done % batch.sr_every:
            raise ValueError("capture at a block boundary")
        self.batch = batch
        side = torch.cuda.Stream(device=batch.device)
        side.wait_stream(torch.cuda.current_stream(batch.device))
        with torch.cuda.stream(side):
            for _ in range(warmup_blocks):
                batch.advance(batch.sr_every)
        torch.cuda.current_stream(batch.device).wait_stream(side)
        torch.cuda.synchronize(batch.device)
        self.graph = torch.cuda.CUDAGraph()
        start = batch.steps_done
        with torch.cuda.graph(self.graph):
            batch.advance(batch.sr_every)
        # capture enqueued nothing: the counter must describe the state, not the recording
        batch.steps_done = start

    def replay(self):
        self.graph.replay()
        self.batch.steps_done += self.batch.sr_every
        return self.batch.block_stats

// Per-walker scalar stages of the step: field shifts (prop/afqmc.py:97-101),
// the phaseless weight update (prop/afqmc.py:116-138) and the block energy
// (prop/blocks.py:59-79).  Memory-bound over (nw, naux) or (nw,) vectors.
#include <cmath>

#include "afqmc_hip.h"
#include "common.h"
#include "device_util.h"

namespace {

// One walker: row w of the (nw, naux) arrays (planes nw * naux apart).
// SPLIT: also the split-bf16 planes of the shifted fields, rows w (real) and
// nw + w (imaginary) of ldk, columns naux..ldk-1 zero (afq::split_bf16).
template <bool SPLIT = false>
__device__ __forceinline__ void field_shift_body(
    int nw, int naux, int w, double sqrt_dt, const double* __restrict__ x, const double* __restrict__ fb,
    const double2* __restrict__ mf, double* __restrict__ xbar, double2* __restrict__ shift_term,
    double2* __restrict__ fb_term, uint16_t* __restrict__ xhi = nullptr, uint16_t* __restrict__ xlo = nullptr,
    long long ldk = 0) {
    __shared__ double sh[16];
    const long long row = (long long)w * naux;
    const long long plane = (long long)nw * naux;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int g = threadIdx.x; g < naux; g += 256) {
        const double fbr = fb[row + g];
        const double fbi = fb[plane + row + g];
        const double2 m = mf[g];
        const double xv = x[row + g];
        // fs = -sqrt_dt * (i fb - mf)
        const double fsr = -sqrt_dt * (-fbi - m.x);
        const double fsi = -sqrt_dt * (fbr - m.y);
        // xbar = x - fs
        const double xr = xv - fsr;
        const double xi = -fsi;
        xbar[row + g] = xr;
        xbar[plane + row + g] = xi;
        if constexpr (SPLIT) {
            const long long er = (long long)w * ldk + g, ei = ((long long)nw + w) * ldk + g;
            afq::split_bf16(xr, xhi[er], xlo[er]);
            afq::split_bf16(xi, xhi[ei], xlo[ei]);
        }
        // shift_term += xbar * mf
        acc[0] += xr * m.x - xi * m.y;
        acc[1] += xr * m.y + xi * m.x;
        // fb_term += x fs - fs^2 / 2
        acc[2] += xv * fsr - 0.5 * (fsr * fsr - fsi * fsi);
        acc[3] += xv * fsi - 0.5 * (2.0 * fsr * fsi);
    }
    if constexpr (SPLIT) {
        for (long long g = naux + threadIdx.x; g < ldk; g += 256) {
            xhi[(long long)w * ldk + g] = 0;
            xlo[(long long)w * ldk + g] = 0;
            xhi[((long long)nw + w) * ldk + g] = 0;
            xlo[((long long)nw + w) * ldk + g] = 0;
        }
    }
    block_sum<256>(acc, sh);
    if (threadIdx.x == 0) {
        shift_term[w] = make_double2(acc[0], acc[1]);
        fb_term[w] = make_double2(acc[2], acc[3]);
    }
}

// One workgroup per walker.
__global__ __launch_bounds__(256) void field_shift_kernel(
    int nw, int naux, double sqrt_dt, const double* __restrict__ x, const double* __restrict__ fb,
    const double2* __restrict__ mf, double* __restrict__ xbar, double2* __restrict__ shift_term,
    double2* __restrict__ fb_term) {
    field_shift_body(nw, naux, blockIdx.x, sqrt_dt, x, fb, mf, xbar, shift_term, fb_term);
}

__global__ __launch_bounds__(256) void field_shift_split_kernel(
    int nw, int naux, double sqrt_dt, const double* __restrict__ x, const double* __restrict__ fb,
    const double2* __restrict__ mf, double* __restrict__ xbar, double2* __restrict__ shift_term,
    double2* __restrict__ fb_term, uint16_t* __restrict__ xhi, uint16_t* __restrict__ xlo, long long ldk) {
    field_shift_body<true>(nw, naux, blockIdx.x, sqrt_dt, x, fb, mf, xbar, shift_term, fb_term, xhi, xlo, ldk);
}

// One workgroup per walker of every system.
__global__ __launch_bounds__(256) void field_shift_grouped_kernel(int count,
                                                                  const afq_system_step* __restrict__ S) {
    const int b = blockIdx.x;
    const int si = afq::find_group(count, b, [&](int i) { return S[i].walker0; });
    const afq_system_step& d = S[si];
    if (d.xbar_hi)  // the VHS operand's split planes too (uniform per workgroup)
        field_shift_body<true>(d.nw, d.naux, b - d.walker0, d.sqrt_dt, d.fields, d.fb, (const double2*)d.mf,
                               d.xbar, (double2*)d.shift_term, (double2*)d.fb_term, d.xbar_hi, d.xbar_lo, d.ldx);
    else
        field_shift_body(d.nw, d.naux, b - d.walker0, d.sqrt_dt, d.fields, d.fb, (const double2*)d.mf,
                         d.xbar, (double2*)d.shift_term, (double2*)d.fb_term);
}

__device__ __forceinline__ void phaseless_body(
    int nw, int nspin, const double2* __restrict__ logdet, const double2* __restrict__ lo_old,
    const double2* __restrict__ st, const double2* __restrict__ ft, const double* __restrict__ w_old,
    const double* __restrict__ pop_old, double e_est, double h0r, double h0i, double dt,
    double sdt, double wfloor, double wcap, double damping, double2* __restrict__ lo_new,
    double* __restrict__ w_new, double* __restrict__ pop_new) {
    __shared__ double sh[64];
    const double pop = *pop_old;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int k = threadIdx.x; k < nw; k += 1024) {
        double2 ld;
        if (nspin == 1) {
            const double2 a = logdet[k];
            ld = make_double2(2.0 * a.x, 2.0 * a.y);
        } else {
            const double2 a = logdet[2 * k];
            const double2 b = logdet[2 * k + 1];
            ld = make_double2(a.x + b.x, a.y + b.y);
        }
        lo_new[k] = ld;
        const double2 lo = lo_old[k];
        const double dlr = ld.x - lo.x;
        const double dli = ld.y - lo.y;
        const double2 s = st[k];
        const double2 f = ft[k];
        // exponent = -sqrt_dt * shift_term + fb_term + dt * (pop_shift + h0_prop)
        const double exr = -sdt * s.x + f.x + dt * (pop + h0r);
        // |exp(exponent) * ratio| and arg(exp(-sqrt_dt * shift_term) * ratio)
        const double mag = exp(exr + dlr);
        const double theta = -sdt * s.y + dli;
        double imp = mag * cos(theta);
        if (imp < wfloor) imp = 0.0;
        if (imp > wcap) imp = 0.0;
        double wn = w_old[k] * imp;
        if (wn > wcap) wn = 0.0;
        w_new[k] = wn;
        acc[0] += wn;
    }
    (void)h0i;
    block_sum<1024>(acc, sh);
    if (threadIdx.x == 0) {
        const double mean = acc[0] / (double)nw;
        const double avg = (mean < 1.0e-300) ? 1.0e-300 : mean;  // NaN propagates
        *pop_new = e_est - damping * (log(avg) / dt);
    }
}

__global__ __launch_bounds__(1024) void phaseless_kernel(
    int nw, int nspin, const double2* __restrict__ logdet, const double2* __restrict__ lo_old,
    const double2* __restrict__ st, const double2* __restrict__ ft, const double* __restrict__ w_old,
    const double* __restrict__ pop_old, double e_est, double h0r, double h0i, double dt,
    double sdt, double wfloor, double wcap, double damping, double2* __restrict__ lo_new,
    double* __restrict__ w_new, double* __restrict__ pop_new) {
    phaseless_body(nw, nspin, logdet, lo_old, st, ft, w_old, pop_old, e_est, h0r, h0i, dt, sdt,
                   wfloor, wcap, damping, lo_new, w_new, pop_new);
}

// One workgroup per system (restricted walkers: log ovlp = 2 log det).
__global__ __launch_bounds__(1024) void phaseless_grouped_kernel(const afq_system_step* __restrict__ S) {
    const afq_system_step& d = S[blockIdx.x];
    phaseless_body(d.nw, 1, (const double2*)d.logdet, (const double2*)d.log_overlaps_in,
                   (const double2*)d.shift_term, (const double2*)d.fb_term, d.weights_in, d.pop_in,
                   d.e_estimate, d.h0_prop, 0.0, d.dt, d.sqrt_dt, d.weight_floor, d.weight_cap,
                   d.damping, (double2*)d.log_overlaps_out, d.weights_out, d.pop_out);
}

__global__ __launch_bounds__(1024) void block_energy_kernel(
    int nw, const double2* __restrict__ e, const double* __restrict__ w, double e_ref,
    double thresh, double alpha, const double* __restrict__ pop, double* __restrict__ out) {
    __shared__ double sh[64];
    block_energy_body(nw, e, w, e_ref, thresh, alpha, pop, out, out + 1, out + 2, sh);
}

}  // namespace

extern "C" int afq_field_shifts(int nw, int naux, double sqrt_dt, const double* fields,
                                const double* fb, const double* mf, double* xbar,
                                double* shift_term, double* fb_term, void* stream) {
    AFQ_REQUIRE(nw > 0 && naux > 0, "afq_field_shifts: bad sizes nw=%d naux=%d", nw, naux);
    AFQ_REQUIRE(fields && fb && mf && xbar && shift_term && fb_term, "afq_field_shifts: null pointer");
    hipLaunchKernelGGL(field_shift_kernel, dim3((unsigned)nw), dim3(256), 0, (hipStream_t)stream, nw,
                       naux, sqrt_dt, fields, fb, (const double2*)mf, xbar, (double2*)shift_term,
                       (double2*)fb_term);
    return afq::check_launch("afq_field_shifts");
}

extern "C" int afq_field_shifts_split(int nw, int naux, double sqrt_dt, const double* fields, const double* fb,
                                      const double* mf, double* xbar, double* shift_term, double* fb_term,
                                      unsigned short* x_hi, unsigned short* x_lo, long long ldk, void* stream) {
    AFQ_REQUIRE(nw > 0 && naux > 0, "afq_field_shifts_split: bad sizes nw=%d naux=%d", nw, naux);
    AFQ_REQUIRE(ldk >= naux, "afq_field_shifts_split: ldk=%lld < naux=%d", ldk, naux);
    AFQ_REQUIRE(fields && fb && mf && xbar && shift_term && fb_term && x_hi && x_lo,
                "afq_field_shifts_split: null pointer");
    hipLaunchKernelGGL(field_shift_split_kernel, dim3((unsigned)nw), dim3(256), 0, (hipStream_t)stream, nw, naux,
                       sqrt_dt, fields, fb, (const double2*)mf, xbar, (double2*)shift_term, (double2*)fb_term,
                       (uint16_t*)x_hi, (uint16_t*)x_lo, ldk);
    return afq::check_launch("afq_field_shifts_split");
}

extern "C" int afq_phaseless_update(int nw, int nspin, const double* logdet_new,
                                    const double* logovlp_old, const double* shift_term,
                                    const double* fb_term, const double* weights_old,
                                    const double* pop_shift_old, double e_estimate,
                                    double h0_prop_re, double h0_prop_im, double dt, double sqrt_dt,
                                    double weight_floor, double weight_cap, double damping,
                                    double* logovlp_new, double* weights_new, double* pop_shift_new,
                                    void* stream) {
    AFQ_REQUIRE(nw > 0 && (nspin == 1 || nspin == 2), "afq_phaseless_update: bad sizes");
    AFQ_REQUIRE(logdet_new && logovlp_old && shift_term && fb_term && weights_old && pop_shift_old &&
                    logovlp_new && weights_new && pop_shift_new,
                "afq_phaseless_update: null pointer");
    AFQ_REQUIRE(dt > 0.0, "afq_phaseless_update: dt must be positive");
    hipLaunchKernelGGL(phaseless_kernel, dim3(1), dim3(1024), 0, (hipStream_t)stream, nw, nspin,
                       (const double2*)logdet_new, (const double2*)logovlp_old,
                       (const double2*)shift_term, (const double2*)fb_term, weights_old,
                       pop_shift_old, e_estimate, h0_prop_re, h0_prop_im, dt, sqrt_dt, weight_floor,
                       weight_cap, damping, (double2*)logovlp_new, weights_new, pop_shift_new);
    return afq::check_launch("afq_phaseless_update");
}

extern "C" int afq_block_energy(int nw, const double* energy, const double* weights, double e_ref,
                                double thresh, double shift_ema, const double* pop_shift,
                                double* out, void* stream) {
    AFQ_REQUIRE(nw > 0, "afq_block_energy: nw=%d", nw);
    AFQ_REQUIRE(energy && weights && pop_shift && out, "afq_block_energy: null pointer");
    hipLaunchKernelGGL(block_energy_kernel, dim3(1), dim3(1024), 0, (hipStream_t)stream, nw,
                       (const double2*)energy, weights, e_ref, thresh, shift_ema, pop_shift, out);
    return afq::check_launch("afq_block_energy");
}

extern "C" int afq_field_shifts_grouped(int count, const afq_system_step* sys, int total_walkers,
                                        void* stream) {
    AFQ_REQUIRE(count > 0 && total_walkers > 0, "afq_field_shifts_grouped: count=%d total_walkers=%d",
                count, total_walkers);
    AFQ_REQUIRE(sys, "afq_field_shifts_grouped: null pointer");
    hipLaunchKernelGGL(field_shift_grouped_kernel, dim3((unsigned)total_walkers), dim3(256), 0,
                       (hipStream_t)stream, count, sys);
    return afq::check_launch("afq_field_shifts_grouped");
}

extern "C" int afq_phaseless_update_grouped(int count, const afq_system_step* sys, void* stream) {
    AFQ_REQUIRE(count > 0, "afq_phaseless_update_grouped: count=%d", count);
    AFQ_REQUIRE(sys, "afq_phaseless_update_grouped: null pointer");
    hipLaunchKernelGGL(phaseless_grouped_kernel, dim3((unsigned)count), dim3(1024), 0,
                       (hipStream_t)stream, sys);
    return afq::check_launch("afq_phaseless_update_grouped");
}

// md_kernels.h -- SCF control, integrator and small bookkeeping kernels.
#pragma once
#include "dev.h"

__global__ void k_set_int(int *p, int v) { *p = v; }
__global__ void k_add_int(int *p, int v) { *p += v; }

__global__ void k_scf_begin(int *ctl, int active)
{
    ctl[C_SCF_ACTIVE] = active;
    ctl[C_ISCF] = 0;
}

struct AspcArgs {
    int n, eo, shift;
    double B[6];
    double W;
};

// one SCF update (pim.c:240-253): iteration 0 takes the ASPC predictor
// (pim.c:382-400) and shifts the history of each ion; later iterations take
// mu = pol^T E on polarisable ions (induced_moment pim.c:617-641); every
// ion also gets its share of u_pol = 1/2 mu.pol^-1.mu (get_upol pim.c:145-161)
__device__ __forceinline__ void cg_block_max3(const double v[3], double *sh, double *wpart, int nb);

// with wpart (256-thread blocks), also the block maxima of |mu| per channel
// (the weight scales of a split-FP16 forward sum over mu).  With begin_ctl
// the kernel opens the SCF itself (k_scf_begin with the loop active): it
// runs ungated and its first thread sets the flag and the iteration count.
__global__ void k_scf_mu(int n, int first, AspcArgs asp, const int *__restrict__ type, const TypeTab *__restrict__ T,
                         double *__restrict__ hist, const double *__restrict__ ef, double *__restrict__ mu,
                         double *__restrict__ upol, const int *active, double *wpart = nullptr,
                         int *begin_ctl = nullptr)
{
    if (begin_ctl) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            begin_ctl[C_SCF_ACTIVE] = 1;
            begin_ctl[C_ISCF] = 0;
        }
    } else if (!*active) {
        return;
    }
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double m[3] = {0.0, 0.0, 0.0};
    if (i < n) {
        int t = type[i];
        const size_t n3 = (size_t)3 * n;
        if (first) {
            for (int a = 0; a < 3; a++) {
                double v = 0.0;
                for (int k = 0; k <= asp.eo; k++) v += asp.B[k] * hist[k * n3 + (size_t)a * n + i];
                m[a] = v;
            }
            for (int k = asp.shift; k > 0; k--)
                for (int a = 0; a < 3; a++) hist[k * n3 + (size_t)a * n + i] = hist[(k - 1) * n3 + (size_t)a * n + i];
        } else if (T->polar[t]) {
            const double *P = T->pol[t];
            double e0 = ef[i], e1 = ef[n + i], e2 = ef[2 * n + i];
            for (int a = 0; a < 3; a++) m[a] = P[0 * 3 + a] * e0 + P[1 * 3 + a] * e1 + P[2 * 3 + a] * e2;
        }
        mu[i] = m[0];
        mu[n + i] = m[1];
        mu[2 * n + i] = m[2];
        double up = 0.0;
        if (T->polar[t]) {
            const double *I = T->invpol[t];
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++) up += m[a] * I[3 * a + b] * m[b];
        }
        upol[i] = 0.5 * up;
    }
    if (wpart) {
        __shared__ double sh[256];
        cg_block_max3(m, sh, wpart, gridDim.x);
    }
}

// ASPC corrector (pim.c:403-420): mu = W pol^T E + (1-W) mu on iteration 0
__global__ void k_scf_correct(int n, AspcArgs asp, const int *__restrict__ type, const TypeTab *__restrict__ T,
                              const double *__restrict__ ef, double *__restrict__ mu, const int *active)
{
    if (!*active) return;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int t = type[i];
    double ms[3] = {mu[i], mu[n + i], mu[2 * n + i]};
    double m[3] = {ms[0], ms[1], ms[2]};
    if (T->polar[t]) {
        const double *P = T->pol[t];
        double e0 = ef[i], e1 = ef[n + i], e2 = ef[2 * n + i];
        for (int a = 0; a < 3; a++) m[a] = P[0 * 3 + a] * e0 + P[1 * 3 + a] * e1 + P[2 * 3 + a] * e2;
    }
    for (int a = 0; a < 3; a++) mu[(size_t)a * n + i] = m[a] * asp.W + (1.0 - asp.W) * ms[a];
}

// get_rmsd_mu pim.c:431-449 per ion (diagonal polarisability only)
__global__ void k_scf_rmsd(int n, const int *__restrict__ type, const TypeTab *__restrict__ T,
                           const double *__restrict__ mu, const double *__restrict__ ef, double *__restrict__ out,
                           const int *active)
{
    if (!*active) return;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int t = type[i];
    double r = 0.0;
    if (T->polar[t]) {
        for (int a = 0; a < 3; a++) {
            double paa = T->pol[t][4 * a];
            if (paa != 0.0) {
                double e = mu[(size_t)a * n + i] / paa - ef[(size_t)a * n + i];
                r += e * e;
            }
        }
    }
    out[i] = r;
}

// one block: get_rmsd_mu pim.c:431-449 summed over the ions (and u_pol
// when upol_i is given), then the loop condition of momentpolaSCF pim.c:236;
// the sums take the fixed order of k_colsum
__global__ void __launch_bounds__(1024) k_scf_check(int n, const int *__restrict__ type,
                                                    const TypeTab *__restrict__ T, const double *__restrict__ mu,
                                                    const double *__restrict__ ef, const double *__restrict__ upol_i,
                                                    int *ctl, double *scal, int npol, int maxit, int minit, double tol)
{
    if (!ctl[C_SCF_ACTIVE]) return;
    __shared__ double sr[1024], su[1024];
    double r = 0.0, u = 0.0;
    for (int i = threadIdx.x; i < n; i += 1024) {
        const int t = type[i];
        if (T->polar[t]) {
            double ri = 0.0;
            for (int c = 0; c < 3; c++) {
                double paa = T->pol[t][4 * c];
                if (paa != 0.0) {
                    double e = mu[(size_t)c * n + i] / paa - ef[(size_t)c * n + i];
                    ri += e * e;
                }
            }
            r += ri;
        }
        if (upol_i) u += upol_i[i];
    }
    sr[threadIdx.x] = r;
    su[threadIdx.x] = u;
    __syncthreads();
    for (int o = 512; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            sr[threadIdx.x] += sr[threadIdx.x + o];
            su[threadIdx.x] += su[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        scal[S_RMSD2] = sr[0];
        if (upol_i) scal[S_UPOL] = su[0];
        int iscf = ctl[C_ISCF] + 1;
        double rmsd = sqrt(sr[0] / (double)(npol > 0 ? npol : 1));
        ctl[C_ISCF] = iscf;
        ctl[C_SCF_ACTIVE] = ((iscf < maxit && rmsd > tol) || iscf < minit) ? 1 : 0;
    }
}

// loop condition of momentpolaSCF pim.c:236
__global__ void k_scf_control(int *ctl, const double *scal, int npol, int maxit, int minit, double tol)
{
    if (!ctl[C_SCF_ACTIVE]) return;
    int iscf = ctl[C_ISCF] + 1;
    double rmsd = sqrt(scal[S_RMSD2] / (double)(npol > 0 ? npol : 1));
    ctl[C_ISCF] = iscf;
    ctl[C_SCF_ACTIVE] = ((iscf < maxit && rmsd > tol) || iscf < minit) ? 1 : 0;
}

__global__ void k_store_hist(int n, const double *mu, double *hist)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 3 * n) hist[i] = mu[i];
}

__global__ void k_axpy3(int m, const double *a, const double *b, double *out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) out[i] = a[i] + b[i];
}

// F = F_short + coul (F_dir + F_rec)  (field.c:352-407, ewald.c:275-288)
__global__ void k_assemble_force(int m, const double *fsr, const double *fdir, const double *frec, double coul,
                                 double *f, int *ctl)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    double v = fsr[i];
    if (coul != 0.0) v += fdir[i] * coul + frec[i] * coul;
    f[i] = v;
    if (!isfinite(v)) ctl[C_NAN] = 1;
}

// ------------------------------------------------------------ integrators
__device__ __forceinline__ double ke_of(double vx, double vy, double vz, double m)
{
    return 0.5 * ((vx * vx + vy * vy + vz * vz) * m);
}

// prop_velocity_verlet integration.c:34-40
__global__ void k_vv_first(int n, double dt, double *r, const double *v, const double *f, double *fold,
                           const double *invm)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double dtsq2 = dt * dt * 0.5;
    for (int a = 0; a < 3; a++) {
        size_t k = (size_t)a * n + i;
        fold[k] = f[k];
        r[k] += v[k] * dt + (f[k] * dtsq2) * invm[i];
    }
}

// prop_velocity_verlet integration.c:49-54
__global__ void k_vv_second(int n, double dt, double *v, const double *f, const double *fold, const double *invm,
                            const double *m, double *ke)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double dt2 = dt * 0.5;
    for (int a = 0; a < 3; a++) {
        size_t k = (size_t)a * n + i;
        v[k] += (fold[k] + f[k]) * dt2 * invm[i];
    }
    ke[i] = ke_of(v[i], v[n + i], v[2 * n + i], m[i]);
}

// prop_pos_vel_verlet integration.c:70-78
__global__ void k_kick_drift(int n, double dt, double *r, double *v, const double *f, const double *invm)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double dt2 = dt * 0.5;
    for (int a = 0; a < 3; a++) {
        size_t k = (size_t)a * n + i;
        v[k] += f[k] * dt2 * invm[i];
    }
    for (int a = 0; a < 3; a++) {
        size_t k = (size_t)a * n + i;
        r[k] += v[k] * dt;
    }
}

// prop_pos_vel_verlet integration.c:83-88
__device__ void sum_to_scalar(double v, double *part, unsigned *cnt, double *out);

// with part (256-thread blocks), the kinetic energy's total also goes to
// *kin (sum_to_scalar)
__global__ void k_kick(int n, double dt, double *v, const double *f, const double *invm, const double *m, double *ke,
                       double *part = nullptr, unsigned *cnt = nullptr, double *kin = nullptr)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double e = 0.0;
    if (i < n) {
        double dt2 = dt * 0.5;
        for (int a = 0; a < 3; a++) {
            size_t k = (size_t)a * n + i;
            v[k] += f[k] * dt2 * invm[i];
        }
        e = ke_of(v[i], v[n + i], v[2 * n + i], m[i]);
        ke[i] = e;
    }
    if (part) sum_to_scalar(e, part, cnt, kin);
}

__global__ void k_ke(int n, const double *v, const double *m, double *ke)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    ke[i] = ke_of(v[i], v[n + i], v[2 * n + i], m[i]);
}

__global__ void k_scale_v(int n, const double *s, double *v, const double *m, double *ke, double *part = nullptr,
                          unsigned *cnt = nullptr, double *kin = nullptr)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double e = 0.0;
    if (i < n) {
        double f = *s;
        v[i] *= f;
        v[n + i] *= f;
        v[2 * n + i] *= f;
        e = ke_of(v[i], v[n + i], v[2 * n + i], m[i]);
        ke[i] = e;
    }
    if (part) sum_to_scalar(e, part, cnt, kin);
}

struct NhcArgs {
    int nh, mults, yo;
    double dt, kT, L;
    double Q[16], iQ[16];   // thermostat masses and their inverses
    double w[7];
};

// chain_nhcn integration.c:189-294 on the thermostat scalars (one thread;
// the forward sweep leaves vxi[n-2] unchanged, integration.c:269-277).  The
// chain is held in registers: loops run over the maximum length 16 with
// constant indices, guarded by the chain length; NH is the register
// capacity (4, 8 or 16) chosen from the chain length at launch so that short
// chains keep every term in registers.
template <int NH>
__global__ void __launch_bounds__(64) k_nhc_chain(NhcArgs a, double *scal, int envt)
{
    const int nh = a.nh;
    double v[NH], x[NH], G[NH];
#pragma unroll
    for (int i = 0; i < NH; i++) {
        v[i] = i < nh ? scal[S_VXI + i] : 0.0;
        x[i] = i < nh ? scal[S_XI + i] : 0.0;
        G[i] = 0.0;
    }
    const double kin = scal[S_KIN];
    double s = 1.0;
    G[0] = 2.0 * kin - a.L * a.kT;
    for (int k = 0; k < a.mults; k++) {
        for (int j = 0; j < a.yo; j++) {
            const double dts = (j < 7 ? a.w[j] : 0.0) * a.dt * (1.0 / (double)a.mults);
            const double dts2 = dts * 0.5, dts4 = dts2 * 0.5, dts8 = dts4 * 0.5;
#pragma unroll
            for (int i = 1; i < NH; i++)
                if (i == nh - 1) {
                    G[i] = (v[i - 1] * v[i - 1] * a.iQ[i - 1] - a.kT);
                    v[i] += G[i] * dts4;
                }
            // the two scalings around each kick use one exponential (same
            // argument, same value)
#pragma unroll
            for (int ih = NH - 2; ih > -1; ih--)
                if (ih <= nh - 2) {
                    const double e = exp(-v[ih + 1] * dts8 * a.iQ[ih + 1]);
                    v[ih] *= e;
                    v[ih] += G[ih] * dts4;
                    v[ih] *= e;
                }
            s *= exp(-v[0] * dts2 * a.iQ[0]);
#pragma unroll
            for (int i = 0; i < NH; i++)
                if (i < nh) x[i] += v[i] * dts2 * a.iQ[i];
            G[0] = 2.0 * kin * s * s - a.L * a.kT;
#pragma unroll
            for (int ih = 0; ih < NH - 2; ih++)
                if (ih < nh - 2) {
                    const double e = exp(-v[ih + 1] * dts8 * a.iQ[ih + 1]);
                    v[ih] *= e;
                    v[ih] += G[ih] * dts4;
                    v[ih] *= e;
                    G[ih + 1] = (v[ih] * v[ih] * a.iQ[ih] - a.kT);
                }
#pragma unroll
            for (int i = 1; i < NH; i++)
                if (i == nh - 1) v[i] += G[i] * dts4;
        }
    }
#pragma unroll
    for (int i = 0; i < NH; i++)
        if (i < nh) {
            scal[S_VXI + i] = v[i];
            scal[S_XI + i] = x[i];
        }
    scal[S_NHC_S] = s;
    if (envt) {
        // the conserved-quantity term of nhcn integration.c:161-176 from the
        // chain just advanced
        double e = a.L * a.kT * x[0] + v[0] * v[0] * 0.5 / a.Q[0];
        double e3 = 0.0, e4 = 0.0;
#pragma unroll
        for (int i = 1; i < NH; i++)
            if (i < nh) {
                e3 += v[i] * v[i] * 0.5 / a.Q[i];
                e4 += a.kT * x[i];
            }
        scal[S_ENVT] = e + e3 + e4;
    }
}

// prop_leap_frog integration.c:101-128
__global__ void k_leapfrog(int n, double dt, double *r, double *rs, double *v, const double *f, const double *invm,
                           const double *m, double *ke)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double dtsq = dt * dt, idt = 0.5 / dt;
    for (int a = 0; a < 3; a++) {
        size_t k = (size_t)a * n + i;
        double u = 2.0 * r[k] - rs[k] + f[k] * dtsq * invm[i];
        v[k] = idt * (u - rs[k]);
        rs[k] = r[k];
        r[k] = u;
    }
    ke[i] = ke_of(v[i], v[n + i], v[2 * n + i], m[i]);
}

// rescale_velocities kinetic.c:87-95; ratio = dt/tauT, or < 0 for NVT
__global__ void k_rescale_lambda(double *scal, int n, double kT, double boltz, double ratio)
{
    double ekin = scal[S_KIN];
    double T = (2.0 * ekin) / (3.0 * (double)n * boltz);
    double lam = ratio >= 0.0 ? sqrt(1.0 + ratio * ((kT / T / boltz) - 1.0)) : sqrt(1.0 + ((kT / T / boltz) - 1.0));
    scal[S_LAMBDA] = lam;
}

__global__ void k_sub_mean(int n, const double *sums, double *v)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 3 * n) return;
    int a = i / n;
    double on = 1.0 / (double)n;
    v[i] -= sums[a] * on;
}

// get_upol pim.c:145-161 per ion
__global__ void k_upol(int n, const int *__restrict__ type, const TypeTab *__restrict__ T,
                       const double *__restrict__ mu, double *__restrict__ out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int t = type[i];
    double up = 0.0;
    if (T->polar[t]) {
        double m[3] = {mu[i], mu[n + i], mu[2 * n + i]};
        const double *I = T->invpol[t];
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) up += m[a] * I[3 * a + b] * m[b];
    }
    out[i] = 0.5 * up;
}

// ---------------------------------------------------------------------------
// Anchored conjugate gradients of the SCF on
//   (pol^-T - T) mu = E_stat      (mu = pol^T E at the fixed point;
// SPD for a symmetric, positive polarisability and a converging SCF)
// with pol^T as preconditioner.  Vectors live on the polarisable rows (ion
// pidx[k]); dot products are per-block partials that every block of the
// next kernel sums in the same fixed order, so the whole iteration stays on
// the device.
#define CG_BLOCK 256

__device__ __forceinline__ double cg_block_sum(double v, double *sh)
{
    // wave sums, then the four wave sums in order (fixed order throughout)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = 0.0;
#pragma unroll
    for (int w = 0; w < CG_BLOCK / 64; w++) r += sh[w];
    __syncthreads();
    return r;
}

// the block's largest |v[a]| of each of three channels into
// wpart[a nb + block] (non-finite values count as zero): the channel maxima
// that scale the weights of a split-FP16 k-space sum over these dipoles
__device__ __forceinline__ void cg_block_max3(const double v[3], double *sh, double *wpart, int nb)
{
    double m[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        m[a] = isfinite(v[a]) ? fabs(v[a]) : 0.0;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) m[a] = fmax(m[a], __shfl_xor(m[a], o));
    }
    if ((threadIdx.x & 63) == 0)
        for (int a = 0; a < 3; a++) sh[(threadIdx.x >> 6) * 3 + a] = m[a];
    __syncthreads();
    if (threadIdx.x < 3) {
        double r = 0.0;
        for (int w = 0; w < CG_BLOCK / 64; w++) r = fmax(r, sh[w * 3 + threadIdx.x]);
        wpart[(size_t)threadIdx.x * nb + blockIdx.x] = r;
    }
    __syncthreads();
}

__device__ __forceinline__ double cg_sum_partials(const double *part, int nb, double *sh)
{
    double v = 0.0;
    for (int b = threadIdx.x; b < nb; b += CG_BLOCK) v += part[b];
    return cg_block_sum(v, sh);
}

// the sum over the grid of one value per thread into *out: block sums in
// part[block] (CG_BLOCK-thread blocks), summed in block order by the block
// that finishes last (atomic count in *cnt, reset by it)
__device__ void sum_to_scalar(double v, double *part, unsigned *cnt, double *out)
{
    __shared__ double sh[CG_BLOCK];
    __shared__ int last;
    const double b = cg_block_sum(v, sh);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = b;
        __threadfence();
        last = atomicAdd(cnt, 1u) == (unsigned)(gridDim.x - 1);
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    const double tot = cg_sum_partials(part, gridDim.x, sh);
    if (threadIdx.x == 0) {
        *cnt = 0u;
        *out = tot;
    }
}

// per ion, (mu/pol - E)^2 summed over the diagonal entries with pol != 0
// (the terms of get_rmsd_mu pim.c:431-449), zero for other ions
__global__ void __launch_bounds__(256) k_resid2(int n, const int *__restrict__ type, const TypeTab *__restrict__ T,
                                                const double *__restrict__ mu, const double *__restrict__ ef,
                                                double *__restrict__ out)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int t = type[i];
    double r = 0.0;
    if (T->polar[t]) {
        for (int c = 0; c < 3; c++) {
            const double paa = T->pol[t][4 * c];
            if (paa != 0.0) {
                const double e = mu[(size_t)c * n + i] / paa - ef[(size_t)c * n + i];
                r += e * e;
            }
        }
    }
    out[i] = r;
}

// the rmsd of get_rmsd_mu (pim.c:431-449) over mu and ef, u_pol (when upol_i
// is given) and the loop condition of momentpolaSCF pim.c:236: per-block
// partial sums over 256 ions each; the block that finishes last (atomic
// count in *cnt, reset by it) sums the partials in block order and applies
// the condition.  The flag read first is written only after every block
// has counted, so all blocks take the same branch.
__global__ void __launch_bounds__(256) k_scf_check(int n, const int *__restrict__ type, const TypeTab *__restrict__ T,
                                                   const double *__restrict__ mu, const double *__restrict__ ef,
                                                   const double *__restrict__ upol_i, int *ctl, double *part,
                                                   unsigned *cnt, double *scal, int npol, int maxit, int minit,
                                                   double tol)
{
    if (!ctl[C_SCF_ACTIVE]) return;
    __shared__ double sh[CG_BLOCK];
    __shared__ int last;
    const int nb = gridDim.x;
    const int i = blockIdx.x * 256 + threadIdx.x;
    double r = 0.0, u = 0.0;
    if (i < n) {
        const int t = type[i];
        if (T->polar[t]) {
            for (int c = 0; c < 3; c++) {
                const double paa = T->pol[t][4 * c];
                if (paa != 0.0) {
                    const double e = mu[(size_t)c * n + i] / paa - ef[(size_t)c * n + i];
                    r += e * e;
                }
            }
        }
        if (upol_i) u = upol_i[i];
    }
    const double sr = cg_block_sum(r, sh);
    const double su = upol_i ? cg_block_sum(u, sh) : 0.0;
    if (threadIdx.x == 0) {
        part[blockIdx.x] = sr;
        part[nb + blockIdx.x] = su;
        __threadfence();
        last = atomicAdd(cnt, 1u) == (unsigned)(nb - 1);
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    const double rs = cg_sum_partials(part, nb, sh);
    const double us = upol_i ? cg_sum_partials(part + nb, nb, sh) : 0.0;
    if (threadIdx.x == 0) {
        *cnt = 0u;
        scal[S_RMSD2] = rs;
        if (upol_i) scal[S_UPOL] = us;
        const int iscf = ctl[C_ISCF] + 1;
        const double rmsd = sqrt(rs / (double)(npol > 0 ? npol : 1));
        ctl[C_ISCF] = iscf;
        ctl[C_SCF_ACTIVE] = ((iscf < maxit && rmsd > tol) || iscf < minit) ? 1 : 0;
    }
}

// the end of an SCF: u_pol of the final dipoles per ion (k_upol, get_upol
// pim.c:145-161) and its total in scal[S_UPOL] (block partials, summed in
// block order by the block that finishes last: atomic count in *cnt, reset
// by it), and dipia_ind[ia][0] = mu (pim.c:281-285, k_store_hist)
__global__ void __launch_bounds__(256) k_upol_hist(int n, const int *__restrict__ type, const TypeTab *__restrict__ T,
                                                   const double *__restrict__ mu, double *__restrict__ upol_i,
                                                   double *__restrict__ hist, double *part, unsigned *cnt,
                                                   double *scal)
{
    __shared__ double sh[CG_BLOCK];
    __shared__ int last;
    const int nb = gridDim.x;
    const int i = blockIdx.x * 256 + threadIdx.x;
    double u = 0.0;
    if (i < n) {
        const int t = type[i];
        double m[3] = {mu[i], mu[n + i], mu[2 * n + i]};
        double up = 0.0;
        if (T->polar[t]) {
            const double *I = T->invpol[t];
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++) up += m[a] * I[3 * a + b] * m[b];
        }
        u = 0.5 * up;
        upol_i[i] = u;
        for (int a = 0; a < 3; a++) hist[(size_t)a * n + i] = m[a];
    }
    const double su = cg_block_sum(u, sh);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = su;
        __threadfence();
        last = atomicAdd(cnt, 1u) == (unsigned)(nb - 1);
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    const double tot = cg_sum_partials(part, nb, sh);
    if (threadIdx.x == 0) {
        *cnt = 0u;
        scal[S_UPOL] = tot;
    }
}

// r = E - pol^-T mu, z = p = pol^T r (= pol^T E - mu), d = 0; partials of
// r.z; with wpart, the block's channel maxima of |p| (cg_block_max3)
__global__ void __launch_bounds__(CG_BLOCK) k_cg_init(int n, int npol, const int *__restrict__ pidx,
                                                      const int *__restrict__ type, const TypeTab *__restrict__ T,
                                                      const double *__restrict__ ef, const double *__restrict__ mu,
                                                      double *r, double *z, double *p, double *d, double *rz_part,
                                                      const int *active, double *wpart = nullptr)
{
    if (!*active) return;
    __shared__ double sh[CG_BLOCK];
    const int k = blockIdx.x * CG_BLOCK + threadIdx.x;
    double acc = 0.0;
    double pv[3] = {0.0, 0.0, 0.0};
    if (k < npol) {
        const int i = pidx[k], t = type[i];
        const double *P = T->pol[t], *I = T->invpol[t];
        double m[3] = {mu[i], mu[n + i], mu[2 * n + i]};
        double rr[3], zz[3];
        for (int a = 0; a < 3; a++) rr[a] = ef[(size_t)a * n + i] - (I[a] * m[0] + I[3 + a] * m[1] + I[6 + a] * m[2]);
        for (int a = 0; a < 3; a++) zz[a] = P[0 * 3 + a] * rr[0] + P[1 * 3 + a] * rr[1] + P[2 * 3 + a] * rr[2];
        for (int a = 0; a < 3; a++) {
            const size_t o = (size_t)a * n + i;
            r[o] = rr[a]; z[o] = zz[a]; p[o] = zz[a]; d[o] = 0.0;
            acc += rr[a] * zz[a];
            pv[a] = zz[a];
        }
    }
    double s = cg_block_sum(acc, sh);
    if (threadIdx.x == 0) rz_part[blockIdx.x] = s;
    if (wpart) cg_block_max3(pv, sh, wpart, gridDim.x);
}

// partials of p.q, q = pol^-T p - Tp (Tp: field of the dipoles p)
__global__ void __launch_bounds__(CG_BLOCK) k_cg_pq(int n, int npol, const int *__restrict__ pidx,
                                                    const int *__restrict__ type, const TypeTab *__restrict__ T,
                                                    const double *__restrict__ p, const double *__restrict__ Tp,
                                                    double *pq_part, const int *active)
{
    if (!*active) return;
    __shared__ double sh[CG_BLOCK];
    const int k = blockIdx.x * CG_BLOCK + threadIdx.x;
    double acc = 0.0;
    if (k < npol) {
        const int i = pidx[k], t = type[i];
        const double *I = T->invpol[t];
        double pv[3] = {p[i], p[n + i], p[2 * n + i]};
        for (int a = 0; a < 3; a++) {
            double q = I[a] * pv[0] + I[3 + a] * pv[1] + I[6 + a] * pv[2] - Tp[(size_t)a * n + i];
            acc += pv[a] * q;
        }
    }
    double s = cg_block_sum(acc, sh);
    if (threadIdx.x == 0) pq_part[blockIdx.x] = s;
}

// mu += d on the polarisable rows
__global__ void k_cg_apply(int n, int npol, const int *__restrict__ pidx, const double *__restrict__ d, double *mu,
                           const int *active)
{
    if (!*active) return;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= npol) return;
    const int i = pidx[k];
    for (int a = 0; a < 3; a++) mu[(size_t)a * n + i] += d[(size_t)a * n + i];
}

// Anchored form of the CG solve: the iterate is the dipoles themselves and
// the field is carried along, E <- E + gamma T p, from one FP64 evaluation
// of the predictor (the low-precision T only acts on the increments).  The
// residual r = E - pol^-T mu and the rmsd of get_rmsd_mu (pim.c:431-449)
// are formed from the updated mu and E on every iteration.
__global__ void __launch_bounds__(CG_BLOCK) k_cga_update(int n, int npol, int nb, int nbq, const int *__restrict__ pidx,
                                                         const int *__restrict__ type, const TypeTab *__restrict__ T,
                                                         const double *__restrict__ p, const double *__restrict__ Tp,
                                                         double *mu, double *ef, double *r, double *z,
                                                         const double *__restrict__ rz_old,
                                                         const double *__restrict__ pq_part, double *rz_new,
                                                         double *rm_part, const int *active)
{
    if (!*active) return;
    __shared__ double sh[CG_BLOCK];
    const int k = blockIdx.x * CG_BLOCK + threadIdx.x;
    // the ion's operands are loaded before the partial sums (the two load
    // chains overlap)
    int i = 0, t = 0;
    double pv[3] = {0, 0, 0}, mv[3] = {0, 0, 0}, ev[3] = {0, 0, 0}, tv[3] = {0, 0, 0};
    if (k < npol) {
        i = pidx[k];
        t = type[i];
        for (int a = 0; a < 3; a++) {
            const size_t o = (size_t)a * n + i;
            pv[a] = p[o]; mv[a] = mu[o]; ev[a] = ef[o]; tv[a] = Tp[o];
        }
    }
    const double rz = cg_sum_partials(rz_old, nb, sh);
    const double pq = cg_sum_partials(pq_part, nbq, sh);
    const double g = (pq != 0.0) ? rz / pq : 0.0;
    double acc = 0.0, rms = 0.0;
    if (k < npol) {
        const double *P = T->pol[t], *I = T->invpol[t];
        double m[3], e[3], rr[3];
        for (int a = 0; a < 3; a++) {
            const size_t o = (size_t)a * n + i;
            m[a] = mv[a] + g * pv[a];
            e[a] = ev[a] + g * tv[a];
            mu[o] = m[a];
            ef[o] = e[a];
        }
        for (int a = 0; a < 3; a++) {
            rr[a] = e[a] - (I[a] * m[0] + I[3 + a] * m[1] + I[6 + a] * m[2]);
            r[(size_t)a * n + i] = rr[a];
            const double paa = P[4 * a];
            if (paa != 0.0) {
                const double d = m[a] / paa - e[a];
                rms += d * d;
            }
        }
        for (int a = 0; a < 3; a++) {
            double zz = P[0 * 3 + a] * rr[0] + P[1 * 3 + a] * rr[1] + P[2 * 3 + a] * rr[2];
            z[(size_t)a * n + i] = zz;
            acc += rr[a] * zz;
        }
    }
    double s = cg_block_sum(acc, sh);
    double sr = cg_block_sum(rms, sh);
    if (threadIdx.x == 0) {
        rz_new[blockIdx.x] = s;
        rm_part[blockIdx.x] = sr;
    }
}

// beta = r.z (new) / r.z (old); p = z + beta p; block 0 records the rmsd
// and applies the loop condition of momentpolaSCF pim.c:236 (one SCF
// iteration per CG step)
__global__ void __launch_bounds__(CG_BLOCK) k_cga_next(int n, int npol, int nb, const int *__restrict__ pidx,
                                                       const double *__restrict__ z, double *p,
                                                       const double *__restrict__ rz_old,
                                                       const double *__restrict__ rz_new,
                                                       const double *__restrict__ rm_part, int *ctl, double *scal,
                                                       int maxit, int minit, double tol, double *wpart = nullptr)
{
    if (!ctl[C_SCF_ACTIVE]) return;
    __shared__ double sh[CG_BLOCK];
    const int k = blockIdx.x * CG_BLOCK + threadIdx.x;
    // the ion's operands before the partial sums (overlapping load chains)
    int i = 0;
    double zv[3] = {0, 0, 0}, po[3] = {0, 0, 0};
    if (k < npol) {
        i = pidx[k];
        for (int a = 0; a < 3; a++) {
            const size_t o = (size_t)a * n + i;
            zv[a] = z[o]; po[a] = p[o];
        }
    }
    const double ro = cg_sum_partials(rz_old, nb, sh);
    const double rn = cg_sum_partials(rz_new, nb, sh);
    const double rm = cg_sum_partials(rm_part, nb, sh);
    const double b = (ro != 0.0) ? rn / ro : 0.0;
    double pv[3] = {0.0, 0.0, 0.0};
    if (k < npol) {
        for (int a = 0; a < 3; a++) {
            pv[a] = zv[a] + b * po[a];
            p[(size_t)a * n + i] = pv[a];
        }
    }
    if (wpart) cg_block_max3(pv, sh, wpart, nb);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        scal[S_RMSD2] = rm;
        int iscf = ctl[C_ISCF] + 1;
        double rmsd = sqrt(rm / (double)(npol > 0 ? npol : 1));
        ctl[C_ISCF] = iscf;
        ctl[C_SCF_ACTIVE] = ((iscf < maxit && rmsd > tol) || iscf < minit) ? 1 : 0;
    }
}

// mu_tot = mu_static + mu_ind (engforce after get_dipoles), with the self
// terms of the force call (q^2 and |mu_tot|^2 per ion, as k_self_terms) and,
// when wpart is given, the per-block maxima of |q|, |mu_tot_a| (four weight
// channels of a split-FP16 force sum) at wpart[c * gridDim.x + block].
// 256-thread blocks.
__global__ void __launch_bounds__(256) k_mu_tot(int n, const double *__restrict__ q, const double *__restrict__ ms,
                                                const double *__restrict__ mi, double *__restrict__ mt,
                                                double *__restrict__ self2n, double *__restrict__ wpart)
{
    __shared__ double sh[4][4];
    const int i = blockIdx.x * 256 + threadIdx.x;
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    if (i < n) {
        const double qi = q[i];
        double m2 = 0.0;
        for (int a = 0; a < 3; a++) {
            const size_t o = (size_t)a * n + i;
            const double m = ms[o] + mi[o];
            mt[o] = m;
            m2 += m * m;
            v[1 + a] = m;
        }
        v[0] = qi;
        self2n[i] = qi * qi;
        self2n[n + i] = m2;
    }
    if (!wpart) return;
#pragma unroll
    for (int c = 0; c < 4; c++) {
        double x = isfinite(v[c]) ? fabs(v[c]) : 0.0;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) x = fmax(x, __shfl_xor(x, o));
        v[c] = x;
    }
    if ((threadIdx.x & 63) == 0)
        for (int c = 0; c < 4; c++) sh[threadIdx.x >> 6][c] = v[c];
    __syncthreads();
    if (threadIdx.x < 4) {
        const double r = fmax(fmax(sh[0][threadIdx.x], sh[1][threadIdx.x]), fmax(sh[2][threadIdx.x], sh[3][threadIdx.x]));
        wpart[(size_t)threadIdx.x * gridDim.x + blockIdx.x] = r;
    }
}

// ---------------------------------------------------------------------------
// scfkO (momentpolaSCFkO pim.c:457-603, induced_moment_inner_kO :647-748),
// per polarisable row.  The inner loop relaxes dmu towards pol^T (f + E(dmu))
// with the real-space dipole field of alpha 0.001 (self term included),
// omega-weighted, while its rmsd exceeds a tenth of the rmsd of f, at most
// 21 sweeps; ctl[C_KO_ACTIVE] gates the sweeps on the device.

// dmu = omega pol^T f; partial sums of f_a^2 (rmsd of zero dipoles in f);
// block 0 opens the inner loop (closed when the SCF is no longer active)
__global__ void __launch_bounds__(CG_BLOCK) k_ko_start(int n, int npol, const int *__restrict__ pidx,
                                                       const int *__restrict__ type, const TypeTab *__restrict__ T,
                                                       const double *__restrict__ f, double *dmu, double omega,
                                                       double *ext_part, int *ctl)
{
    const bool on = ctl[C_SCF_ACTIVE] != 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ctl[C_KO_ACTIVE] = on ? 1 : 0;
        ctl[C_KO_ITER] = 0;
    }
    if (!on) return;
    __shared__ double sh[CG_BLOCK];
    const int k = blockIdx.x * CG_BLOCK + threadIdx.x;
    double acc = 0.0;
    if (k < npol) {
        const int i = pidx[k], t = type[i];
        const double *P = T->pol[t];
        const double f0 = f[i], f1 = f[n + i], f2 = f[2 * n + i];
        const double fv[3] = {f0, f1, f2};
        for (int a = 0; a < 3; a++)
            if (P[4 * a] != 0.0) acc += fv[a] * fv[a];
        for (int a = 0; a < 3; a++) {
            const double m = P[0 * 3 + a] * f0 + P[1 * 3 + a] * f1 + P[2 * 3 + a] * f2;
            dmu[(size_t)a * n + i] = m * omega;
        }
    }
    const double s = cg_block_sum(acc, sh);
    if (threadIdx.x == 0) ext_part[blockIdx.x] = s;
}

// one sweep: E = f + (sself dmu + E_dir); rmsd partials of (dmu, E);
// dmu = (1 - omega) dmu + omega pol^T E
__global__ void __launch_bounds__(CG_BLOCK) k_ko_sweep(int n, int npol, const int *__restrict__ pidx,
                                                       const int *__restrict__ type, const TypeTab *__restrict__ T,
                                                       const double *__restrict__ f, const double *__restrict__ efdir,
                                                       double sself, double omega, double *dmu, double *in_part,
                                                       const int *ctl)
{
    if (!ctl[C_KO_ACTIVE]) return;
    __shared__ double sh[CG_BLOCK];
    const int k = blockIdx.x * CG_BLOCK + threadIdx.x;
    double acc = 0.0;
    if (k < npol) {
        const int i = pidx[k], t = type[i];
        const double *P = T->pol[t];
        double m[3], e[3];
        for (int a = 0; a < 3; a++) {
            const size_t o = (size_t)a * n + i;
            m[a] = dmu[o];
            e[a] = f[o] + (sself * m[a] + efdir[o]);
        }
        for (int a = 0; a < 3; a++) {
            const double paa = P[4 * a];
            if (paa != 0.0) {
                const double d = m[a] / paa - e[a];
                acc += d * d;
            }
        }
        for (int a = 0; a < 3; a++) {
            const double nm = P[0 * 3 + a] * e[0] + P[1 * 3 + a] * e[1] + P[2 * 3 + a] * e[2];
            dmu[(size_t)a * n + i] = (1.0 - omega) * m[a] + omega * nm;
        }
    }
    const double s = cg_block_sum(acc, sh);
    if (threadIdx.x == 0) in_part[blockIdx.x] = s;
}

// loop condition of the inner loop (pim.c:698): another sweep while
// sweeps <= 20 and rmsd > 0.1 rmsd(f); one block
__global__ void __launch_bounds__(CG_BLOCK) k_ko_check(int nb, int npol, const double *__restrict__ in_part,
                                                       const double *__restrict__ ext_part, int *ctl)
{
    if (!ctl[C_KO_ACTIVE]) return;
    __shared__ double sh[CG_BLOCK];
    const double rin = cg_sum_partials(in_part, nb, sh);
    const double rext = cg_sum_partials(ext_part, nb, sh);
    if (threadIdx.x == 0) {
        const int it = ctl[C_KO_ITER] + 1;
        ctl[C_KO_ITER] = it;
        ctl[C_KO_TOTAL] += 1;
        const double np = (double)(npol > 0 ? npol : 1);
        const double ri = sqrt(rin / np), re = sqrt(rext / np);
        ctl[C_KO_ACTIVE] = (it <= 20 && ri > 0.1 * re) ? 1 : 0;
    }
}

// residual field of the next update (pim.c:579-586): f = E - mu diag(pol^-1)
__global__ void k_ko_resid(int n, int npol, const int *__restrict__ pidx, const int *__restrict__ type,
                           const TypeTab *__restrict__ T, const double *__restrict__ ef,
                           const double *__restrict__ mu, double *f, const int *active)
{
    if (!*active) return;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= npol) return;
    const int i = pidx[k], t = type[i];
    const double *I = T->invpol[t];
    for (int a = 0; a < 3; a++) {
        const size_t o = (size_t)a * n + i;
        f[o] = ef[o] - mu[o] * I[4 * a];
    }
}


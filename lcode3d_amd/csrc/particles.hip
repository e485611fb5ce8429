// Particle kernels: field-free estimate, fused interpolation + pusher, and
// coarse->fine interpolation + TSC deposition.  FP64 throughout.

#include "common.hpp"

namespace lcode {

namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ int clampi(int v, int n) {
    return v < 0 ? 0 : (v >= n ? n - 1 : v);
}

// Field-free half-step estimate with coordinate-only reflection using >= and
// <= (lcode.py:369-391).  Returns the estimated offsets.
__device__ __forceinline__ void estimate_one(double m, double xi0, double yi0,
                                             double xo, double yo, double px,
                                             double py, double pz, double dxi,
                                             double r, double& ex,
                                             double& ey) {
    double x = xi0 + xo, y = yi0 + yo;
    const double gamma_m = sqrt(m * m + pz * pz + px * px + py * py);
    x += px / (gamma_m - pz) * dxi;
    y += py / (gamma_m - pz) * dxi;
    if (x >= +r) x = +2 * r - x;
    if (x <= -r) x = -2 * r - x;
    if (y >= +r) y = +2 * r - y;
    if (y <= -r) y = -2 * r - y;
    ex = x - xi0;
    ey = y - yi0;
}

__global__ void __launch_bounds__(kBlock)
estimate_kernel(EstimateArgs a) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.n) return;
    double ex, ey;
    estimate_one(a.m[k], a.x_init[k], a.y_init[k], a.prev_x_offt[k],
                 a.prev_y_offt[k], a.px[k], a.py[k], a.pz[k], a.dxi,
                 a.reflect, ex, ey);
    a.out_x[k] = ex;
    a.out_y[k] = ey;
}

// move_smart_kernel (lcode.py:704-793): midpoint position, 6-field TSC
// gather, two fixed-point iterations of the half-step momentum, offset and
// momentum update, reflection with strict > / < flipping the momentum.
__global__ void __launch_bounds__(kBlock)
push_kernel(PushArgs a) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.n) return;
    const double m = a.m[k], q = a.q[k];
    const double opx = a.px[k], opy = a.py[k], opz = a.pz[k];
    const double xi0 = a.x_init[k], yi0 = a.y_init[k];
    double x_offt = a.prev_x_offt[k], y_offt = a.prev_y_offt[k];
    double ex, ey;
    if (a.est_x) {
        ex = a.est_x[k];
        ey = a.est_y[k];
    } else {
        estimate_one(m, xi0, yi0, x_offt, y_offt, opx, opy, opz, a.dxi,
                     a.reflect, ex, ey);
    }
    const double x_half = xi0 + (x_offt + ex) / 2;
    const double y_half = yi0 + (y_offt + ey) / 2;
    Tsc t;
    tsc_weights(x_half, y_half, a.N, a.h, t, a.fault);
    const double Ex = interp9(a.F[0], a.N, t);
    const double Ey = interp9(a.F[1], a.N, t);
    const double Ez = interp9(a.F[2], a.N, t);
    const double Bx = interp9(a.F[3], a.N, t);
    const double By = interp9(a.F[4], a.N, t);
    const double Bz = interp9(a.F[5], a.N, t);

    double px = opx, py = opy, pz = opz;
    double dpx = 0, dpy = 0, dpz = 0;
#pragma unroll
    for (int it = 0; it < 2; ++it) {
        const double gamma_m = sqrt(m * m + pz * pz + px * px + py * py);
        const double vx = px / gamma_m, vy = py / gamma_m, vz = pz / gamma_m;
        const double factor_1 = q * a.dxi / (1 - pz / gamma_m);
        dpx = factor_1 * (Ex + vy * Bz - vz * By);
        dpy = factor_1 * (Ey - vx * Bz + vz * Bx);
        dpz = factor_1 * (Ez + vx * By - vy * Bx);
        px = opx + dpx / 2;
        py = opy + dpy / 2;
        pz = opz + dpz / 2;
    }
    const double gamma_m = sqrt(m * m + pz * pz + px * px + py * py);
    x_offt += px / (gamma_m - pz) * a.dxi;
    y_offt += py / (gamma_m - pz) * a.dxi;
    px = opx + dpx;
    py = opy + dpy;
    pz = opz + dpz;

    const double r = a.reflect;
    double x = xi0 + x_offt, y = yi0 + y_offt;
    if (x > +r) { x = +2 * r - x; x_offt = x - xi0; px = -px; }
    if (x < -r) { x = -2 * r - x; x_offt = x - xi0; px = -px; }
    if (y > +r) { y = +2 * r - y; y_offt = y - yi0; py = -py; }
    if (y < -r) { y = -2 * r - y; y_offt = y - yi0; py = -py; }

    a.out_x[k] = x_offt;
    a.out_y[k] = y_offt;
    a.out_px[k] = px;
    a.out_py[k] = py;
    a.out_pz[k] = pz;
}

__global__ void __launch_bounds__(kBlock)
zero4_kernel(double* a, double* b, double* c, double* d, int n) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    a[k] = 0; b[k] = 0; c[k] = 0; d[k] = 0;
}

__global__ void __launch_bounds__(kBlock)
add_kernel(double* __restrict__ a, const double* __restrict__ b, int n) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) a[k] += b[k];
}

// deposit_kernel (lcode.py:626-663): one fine particle per thread,
// bilinear coarse->fine interpolation (lcode.py:591-621), then 4 x 9 FP64
// atomic adds into the grids.
__global__ void __launch_bounds__(kBlock)
deposit_atomic_kernel(DepositArgs a) {
    const int fk = blockIdx.x * blockDim.x + threadIdx.x;
    if (fk >= a.nf * a.nf) return;
    const int fi = fk / a.nf, fj = fk % a.nf;
    const double ipi = a.ip[fi], ini = a.in[fi];
    const double ipj = a.ip[fj], inj = a.in[fj];
    const double A = ipi * ipj, B = ipi * inj, C = ini * ipj, D = ini * inj;
    const int nc = a.nc;
    // neighbour indices come from the caller: clamp so that bad ones cannot
    // read outside the coarse arrays
    const int pi = clampi((int)a.P[fi], nc), ni = clampi((int)a.Nn[fi], nc);
    const int pj = clampi((int)a.P[fj], nc), nj = clampi((int)a.Nn[fj], nc);
    const int pp = pi * nc + pj, pn = pi * nc + nj;
    const int np = ni * nc + pj, nn = ni * nc + nj;
#define LCODE_MIX(arr) (A * arr[pp] + B * arr[pn] + C * arr[np] + D * arr[nn])
    const double x = a.fg[fi] + LCODE_MIX(a.x_offt);
    const double y = a.fg[fj] + LCODE_MIX(a.y_offt);
    const double s = a.smallness;
    const double m = s * LCODE_MIX(a.m);
    const double q = s * LCODE_MIX(a.q);
    const double px = s * LCODE_MIX(a.px);
    const double py = s * LCODE_MIX(a.py);
    const double pz = s * LCODE_MIX(a.pz);
#undef LCODE_MIX
    const double gamma_m = sqrt(m * m + px * px + py * py + pz * pz);
    const double dro = q / (1 - pz / gamma_m);
    const double djx = px * (dro / gamma_m);
    const double djy = py * (dro / gamma_m);
    const double djz = pz * (dro / gamma_m);
    Tsc t;
    tsc_weights(x, y, a.N, a.h, t, a.fault);
    const int N = a.N;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const size_t idx = (size_t)(t.i + stencil_di(k)) * N + (t.j + stencil_dj(k));
        unsafeAtomicAdd(a.ro + idx, dro * t.w[k]);
        unsafeAtomicAdd(a.jx + idx, djx * t.w[k]);
        unsafeAtomicAdd(a.jy + idx, djy * t.w[k]);
        unsafeAtomicAdd(a.jz + idx, djz * t.w[k]);
    }
}

inline int blocks_for(long n) { return (int)((n + kBlock - 1) / kBlock); }

}  // namespace

int launch_estimate(const EstimateArgs& a, hipStream_t s) {
    if (a.n == 0) return LCODE_OK;
    estimate_kernel<<<blocks_for(a.n), kBlock, 0, s>>>(a);
    LCODE_HIP(hipGetLastError());
    return LCODE_OK;
}

int launch_push(const PushArgs& a, hipStream_t s) {
    if (a.n == 0) return LCODE_OK;
    push_kernel<<<blocks_for(a.n), kBlock, 0, s>>>(a);
    LCODE_HIP(hipGetLastError());
    return LCODE_OK;
}

int launch_deposit(const DepositArgs& a, hipStream_t s) {
    const int n2 = a.N * a.N;
    zero4_kernel<<<blocks_for(n2), kBlock, 0, s>>>(a.ro, a.jx, a.jy, a.jz, n2);
    LCODE_HIP(hipGetLastError());
    const long nfp = (long)a.nf * a.nf;
    if (nfp > 0) {
        deposit_atomic_kernel<<<blocks_for(nfp), kBlock, 0, s>>>(a);
        LCODE_HIP(hipGetLastError());
    }
    // The ion background is added last "to preserve more float precision"
    // (lcode.py:686-687).
    if (a.ro_initial) {
        add_kernel<<<blocks_for(n2), kBlock, 0, s>>>(a.ro, a.ro_initial, n2);
        LCODE_HIP(hipGetLastError());
    }
    return LCODE_OK;
}

}  // namespace lcode

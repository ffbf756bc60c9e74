// nova_fft.hip -- FFT Poisson solve and projection on gfx950.
//
// Replaces poisson::rhs/solve (src/poisson/poisson.cpp:167-181,223-265),
// which gathers the field into an i-fastest array and runs two full complex
// 3-D FFTs through heFFTe/FFTW.  Here the field stays k-fastest and the
// transform uses the Hermitian half along k:
//
//   pass 1  k_fwd_z    source (div(F)/s fused, or a scalar rhs) -> r2c
//                      along k, two real lines packed per complex FFT;
//                      writes H[kc][i][j] (j fastest)
//   pass 2  k_fft_y    c2c along j (contiguous lines); for a walled y
//                      the even extension of length 2 N_y (DCT-I)
//   pass 3  k_fft_x    c2c along i (batch along j), multiply by -1/k'^2
//                      (0 for k'^2 < 1e-16, poisson.cpp:237-244), inverse
//                      c2c along i, all in one LDS tile
//   pass 4  k_fft_y    inverse along j
//   pass 5  k_inv_z    c2r along k, p = v / (N_x N_y N_z) (poisson.cpp:259)
//
// Each pass reads and writes the half spectrum once (16 B per complex
// value).  Line FFTs run in LDS as mixed-radix 4/2/3 passes: DIT from a
// digit-reversed load, or DIF into a digit-reversed store.
#include "nova_internal.hpp"

#include <cmath>

struct nova_poisson;

namespace nova {

struct FftLen {
    int n = 0;
    int nr = 0;
    int radix[16] = {0};
    double2 *tw = nullptr;  // W_n^e = exp(-2 pi i e / n), e in [0, n)
    int *perm = nullptr;    // DIT load position of element e
};

struct FftDev {
    int n, nr;
    int radix[16];
};

}  // namespace nova

struct nova_poisson {
    nova_ctx *ctx = nullptr;
    int Nx = 0, Ny = 0, Nz = 0;
    int sy = 0;      // stored y extent
    int wall_y = 0;
    int Ty = 0;      // y transform length
    int Kc = 0;      // Nz/2 + 1
    nova::FftLen fx, fy, fz;
    double2 *H = nullptr;
    double *kpx = nullptr, *kpy = nullptr, *kpz = nullptr;
    double norm = 1.0;
    double *ptmp = nullptr;  // potential when the caller passes none
};

namespace nova {

namespace {

constexpr int kFftThreads = 256;
constexpr double kEps = 1.0e-16;  // src/Tools/Tools.cpp:7

__device__ inline double2 cmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ inline double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
// multiply by -i (forward) or +i (inverse)
__device__ inline double2 rot(double2 a, bool inv) { return inv ? make_double2(-a.y, a.x) : make_double2(a.y, -a.x); }

__device__ inline void dft_small(double2 *a, int R, bool inv)
{
    if (R == 2) {
        double2 t = a[0];
        a[0] = cadd(t, a[1]);
        a[1] = csub(t, a[1]);
    } else if (R == 4) {
        double2 s02 = cadd(a[0], a[2]), d02 = csub(a[0], a[2]);
        double2 s13 = cadd(a[1], a[3]), d13 = rot(csub(a[1], a[3]), inv);
        a[0] = cadd(s02, s13);
        a[2] = csub(s02, s13);
        a[1] = cadd(d02, d13);
        a[3] = csub(d02, d13);
    } else if (R == 3) {
        const double c = -0.5, s = 0.86602540378443864676;
        double2 sum = cadd(a[1], a[2]);
        double2 dif = csub(a[1], a[2]);
        double2 t = make_double2(a[0].x + c * sum.x, a[0].y + c * sum.y);
        double2 r = rot(make_double2(s * dif.x, s * dif.y), inv);  // -i s (a1-a2) forward
        a[0] = cadd(a[0], sum);
        a[1] = cadd(t, r);
        a[2] = csub(t, r);
    }
}

// In-place line FFTs in LDS.  Element e of line l sits at
// buf[l*lstride + e*estride].  DIT expects digit-reversed input and leaves
// natural order; DIF takes natural order and leaves digit-reversed output.
// batch_fast picks which index consecutive threads walk (the one that is
// contiguous in LDS).
__device__ void lds_fft(double2 *buf, int nlines, int lstride, int estride, const FftDev &P,
                        const double2 *tw, bool inv, bool dit, bool batch_fast)
{
    const int n = P.n;
    int Lps[16];
    {
        int acc = 1;
        for (int s = 0; s < P.nr; s++) { Lps[s] = acc; acc *= P.radix[s]; }
    }
    for (int it = 0; it < P.nr; it++) {
        const int s = dit ? it : P.nr - 1 - it;
        const int R = P.radix[s];
        const int Lp = Lps[s];
        const int L = Lp * R;
        const int nb = n / R;
        const int tstep = n / L;
        const int total = nb * nlines;
        for (int w = threadIdx.x; w < total; w += blockDim.x) {
            int line, bf;
            if (batch_fast) { line = w % nlines; bf = w / nlines; }
            else { line = w / nb; bf = w - line * nb; }
            const int g = bf / Lp;
            const int q = bf - g * Lp;
            double2 *b0 = buf + line * lstride + (g * L + q) * estride;
            const int st = Lp * estride;
            double2 a[4];
            for (int p = 0; p < R; p++) a[p] = b0[p * st];
            if (dit) {
                for (int p = 1; p < R; p++) {
                    double2 wv = tw[p * q * tstep];
                    if (inv) wv.y = -wv.y;
                    a[p] = cmul(a[p], wv);
                }
                dft_small(a, R, inv);
            } else {
                dft_small(a, R, inv);
                for (int p = 1; p < R; p++) {
                    double2 wv = tw[p * q * tstep];
                    if (inv) wv.y = -wv.y;
                    a[p] = cmul(a[p], wv);
                }
            }
            for (int p = 0; p < R; p++) b0[p * st] = a[p];
        }
        __syncthreads();
    }
}

__device__ inline void load_tables(double2 *stw, int *sperm, const double2 *tw, const int *perm, int n)
{
    for (int e = threadIdx.x; e < n; e += blockDim.x) {
        stw[e] = tw[e];
        if (sperm) sperm[e] = perm[e];
    }
    __syncthreads();
}

__device__ inline int64_t lin(const Grid &g, int i, int j, int k)
{
    return ((int64_t)i * g.s[1] + j) * g.s[2] + k;
}

__device__ inline double ldg(const double *__restrict__ f, const Grid &g, int nc, int c, int i, int j, int k)
{
    int fi, fj, fk;
    int ii = bmap(i, g.n[0], g.wall[0], fi);
    int jj = bmap(j, g.n[1], g.wall[1], fj);
    int kk = bmap(k, g.n[2], g.wall[2], fk);
    double s = bsign(nc, c, 0, fi) * bsign(nc, c, 1, fj) * bsign(nc, c, 2, fk);
    return s * f[c * g.np + lin(g, ii, jj, kk)];
}

// div(F) * scale at a stored node (src/ex/div/div.hxx:37-77).
__device__ inline double div_at(const Grid &g, const double *__restrict__ F, int i, int j, int k, double scale)
{
    double acc = 0.0;
    double ax = 0.0, ay = 0.0, az = 0.0;
    for (int l = 0; l < g.h; l++) {
        ax += g.a[0][l] * (ldg(F, g, 3, 0, i + 1 + l, j, k) - ldg(F, g, 3, 0, i - 1 - l, j, k));
        ay += g.a[1][l] * (ldg(F, g, 3, 1, i, j + 1 + l, k) - ldg(F, g, 3, 1, i, j - 1 - l, k));
        az += g.a[2][l] * (ldg(F, g, 3, 2, i, j, k + 1 + l) - ldg(F, g, 3, 2, i, j, k - 1 - l));
    }
    acc = ax + ay + az * g.zinv[k];
    return acc * scale;
}

constexpr int kZLines = 16;  // real lines per block in the k passes

// pass 1: source -> half spectrum H[kc][i][j].
__global__ void __launch_bounds__(kFftThreads) k_fwd_z(Grid g, FftDev P, const double2 *__restrict__ tw,
                                                       const int *__restrict__ perm, const double *__restrict__ src,
                                                       int src_is_vector, double scale, double2 *__restrict__ H,
                                                       int Nx, int sy)
{
    extern __shared__ double2 smem[];
    const int n = P.n;  // Nz
    double2 *stw = smem;
    int *sperm = (int *)(smem + n);
    double2 *buf = smem + n + (n + 1) / 2;  // 8 complex lines
    load_tables(stw, sperm, tw, perm, n);
    const int i = blockIdx.y;
    const int j0 = blockIdx.x * kZLines;
    for (int e = threadIdx.x; e < kZLines * n; e += blockDim.x) {
        const int jj = e / n, k = e - jj * n;
        const int j = j0 + jj;
        double v = 0.0;
        if (j < sy) v = src_is_vector ? div_at(g, src, i, j, k, scale) : src[lin(g, i, j, k)] * scale;
        double *slot = (double *)&buf[(jj >> 1) * n + sperm[k]];
        slot[jj & 1] = v;
    }
    __syncthreads();
    lds_fft(buf, kZLines / 2, n, 1, P, stw, false, true, false);
    const int Kc = n / 2 + 1;
    for (int e = threadIdx.x; e < kZLines * Kc; e += blockDim.x) {
        const int jj = e % kZLines, kc = e / kZLines;
        const int j = j0 + jj;
        if (j >= sy) continue;
        const double2 *Z = buf + (jj >> 1) * n;
        const double2 z = Z[kc], zm = Z[(n - kc) % n];
        double2 out;
        if ((jj & 1) == 0) out = make_double2(0.5 * (z.x + zm.x), 0.5 * (z.y - zm.y));
        else out = make_double2(0.5 * (z.y + zm.y), -0.5 * (z.x - zm.x));
        H[((int64_t)kc * Nx + i) * sy + j] = out;
    }
}

// passes 2 and 4: lines along j (contiguous).  `ext` = even extension of
// the sy = N_y + 1 stored nodes to length 2 N_y (DCT-I through the FFT).
__global__ void __launch_bounds__(kFftThreads) k_fft_y(FftDev P, const double2 *__restrict__ tw,
                                                       const int *__restrict__ perm, double2 *__restrict__ H,
                                                       int64_t nlines_total, int sy, int lb, int ext, int inv)
{
    extern __shared__ double2 smem[];
    const int n = P.n;
    double2 *stw = smem;
    int *sperm = (int *)(smem + n);
    double2 *buf = smem + n + (n + 1) / 2;
    load_tables(stw, sperm, tw, perm, n);
    const int64_t l0 = (int64_t)blockIdx.x * lb;
    const int Ny = ext ? sy - 1 : sy;
    for (int e = threadIdx.x; e < lb * n; e += blockDim.x) {
        const int ln = e / n, q = e - ln * n;
        const int64_t L = l0 + ln;
        double2 v = make_double2(0.0, 0.0);
        if (L < nlines_total) {
            const int src = (ext && q > Ny) ? 2 * Ny - q : q;
            v = H[L * sy + src];
        }
        buf[ln * n + sperm[q]] = v;
    }
    __syncthreads();
    lds_fft(buf, lb, n, 1, P, stw, inv != 0, true, false);
    for (int e = threadIdx.x; e < lb * sy; e += blockDim.x) {
        const int ln = e / sy, q = e - ln * sy;
        const int64_t L = l0 + ln;
        if (L < nlines_total) H[L * sy + q] = buf[ln * n + q];
    }
}

// pass 3: lines along i, batch of jb consecutive j; forward, divide by
// -k'^2, inverse.
__global__ void __launch_bounds__(kFftThreads) k_fft_x(FftDev P, const double2 *__restrict__ tw,
                                                       const int *__restrict__ perm, double2 *__restrict__ H,
                                                       int Nx, int sy, int jb, const double *__restrict__ kpx,
                                                       const double *__restrict__ kpy,
                                                       const double *__restrict__ kpz)
{
    extern __shared__ double2 smem[];
    const int n = P.n;  // Nx
    double2 *stw = smem;
    int *sperm = (int *)(smem + n);
    double2 *buf = smem + n + (n + 1) / 2;
    load_tables(stw, sperm, tw, perm, n);
    const int kc = blockIdx.y;
    const int j0 = blockIdx.x * jb;
    const double2 *Hs = H + (int64_t)kc * Nx * sy;
    for (int e = threadIdx.x; e < n * jb; e += blockDim.x) {
        const int jj = e % jb, i = e / jb;
        const int j = j0 + jj;
        double2 v = make_double2(0.0, 0.0);
        if (j < sy) v = Hs[(int64_t)i * sy + j];
        buf[sperm[i] * jb + jj] = v;
    }
    __syncthreads();
    lds_fft(buf, jb, 1, jb, P, stw, false, true, true);
    const double kz = kpz[kc];
    for (int e = threadIdx.x; e < n * jb; e += blockDim.x) {
        const int jj = e % jb, i = e / jb;
        const int j = j0 + jj;
        if (j >= sy) continue;
        double k2;
        {
#pragma clang fp contract(off)
            const double kx = kpx[i], ky = kpy[j];
            k2 = kx * kx + ky * ky + kz * kz;
        }
        double2 v = buf[i * jb + jj];
        if (k2 < kEps) v = make_double2(0.0, 0.0);
        else v = make_double2(-v.x / k2, -v.y / k2);
        buf[i * jb + jj] = v;
    }
    __syncthreads();
    lds_fft(buf, jb, 1, jb, P, stw, true, false, true);
    double2 *Hd = H + (int64_t)kc * Nx * sy;
    for (int e = threadIdx.x; e < n * jb; e += blockDim.x) {
        const int jj = e % jb, i = e / jb;
        const int j = j0 + jj;
        if (j < sy) Hd[(int64_t)i * sy + j] = buf[sperm[i] * jb + jj];
    }
}

// pass 5: half spectrum -> real lines along k, p = v / norm.
__global__ void __launch_bounds__(kFftThreads) k_inv_z(Grid g, FftDev P, const double2 *__restrict__ tw,
                                                       const int *__restrict__ perm,
                                                       const double2 *__restrict__ H, double *__restrict__ p,
                                                       int Nx, int sy, double inv_norm)
{
    extern __shared__ double2 smem[];
    const int n = P.n;
    double2 *stw = smem;
    int *sperm = (int *)(smem + n);
    double2 *buf = smem + n + (n + 1) / 2;
    load_tables(stw, sperm, tw, perm, n);
    const int i = blockIdx.y;
    const int j0 = blockIdx.x * kZLines;
    const int Kc = n / 2 + 1;
    // even j: A terms; odd j: i*B terms (Z = A + iB, Hermitian-extended)
    for (int phase = 0; phase < 2; phase++) {
        for (int e = threadIdx.x; e < (kZLines / 2) * Kc; e += blockDim.x) {
            const int pr = e % (kZLines / 2), kc = e / (kZLines / 2);
            const int j = j0 + 2 * pr + phase;
            double2 v = make_double2(0.0, 0.0);
            if (j < sy) v = H[((int64_t)kc * Nx + i) * sy + j];
            double2 *Z = buf + pr * n;
            const bool mirror = kc > 0 && 2 * kc < n;
            if (phase == 0) {
                Z[sperm[kc]] = v;
                if (mirror) Z[sperm[n - kc]] = make_double2(v.x, -v.y);
            } else {
                double2 t = Z[sperm[kc]];
                Z[sperm[kc]] = make_double2(t.x - v.y, t.y + v.x);
                if (mirror) {
                    double2 u = Z[sperm[n - kc]];  // + i * conj(v)
                    Z[sperm[n - kc]] = make_double2(u.x + v.y, u.y + v.x);
                }
            }
        }
        __syncthreads();
    }
    lds_fft(buf, kZLines / 2, n, 1, P, stw, true, true, false);
    for (int e = threadIdx.x; e < kZLines * n; e += blockDim.x) {
        const int jj = e / n, k = e - jj * n;
        const int j = j0 + jj;
        if (j >= sy) continue;
        const double2 z = buf[(jj >> 1) * n + k];
        p[lin(g, i, j, k)] = ((jj & 1) ? z.y : z.x) * inv_norm;
    }
}

// F <- F - s * grad(p) (app/TG_Piosson/RKLoop.H:36, app/OT_Poisson/BEqn.H:40).
__global__ void k_correct(Grid g, const double *__restrict__ p, double s, double *__restrict__ F)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int i = blockIdx.z;
    if (k >= g.s[2] || j >= g.s[1]) return;
    double gx = 0.0, gy = 0.0, gz = 0.0;
    for (int l = 0; l < g.h; l++) {
        gx += g.a[0][l] * (ldg(p, g, 1, 0, i + 1 + l, j, k) - ldg(p, g, 1, 0, i - 1 - l, j, k));
        gy += g.a[1][l] * (ldg(p, g, 1, 0, i, j + 1 + l, k) - ldg(p, g, 1, 0, i, j - 1 - l, k));
        gz += g.a[2][l] * (ldg(p, g, 1, 0, i, j, k + 1 + l) - ldg(p, g, 1, 0, i, j, k - 1 - l));
    }
    gz *= g.zinv[k];
    const int64_t q = lin(g, i, j, k);
    F[q] = F[q] - s * gx;
    F[g.np + q] = F[g.np + q] - s * gy;
    F[2 * g.np + q] = F[2 * g.np + q] - s * gz;
}

// ---- host side ----

void plan_radices(int n, FftLen &f)
{
    f.n = n;
    f.nr = 0;
    int r = n;
    while (r % 4 == 0) { f.radix[f.nr++] = 4; r /= 4; }
    while (r % 2 == 0) { f.radix[f.nr++] = 2; r /= 2; }
    while (r % 3 == 0) { f.radix[f.nr++] = 3; r /= 3; }
}

int make_len(FftLen &f, int n)
{
    plan_radices(n, f);
    int prod = 1;
    for (int s = 0; s < f.nr; s++) prod *= f.radix[s];
    if (prod != n || f.nr > 16) return fail(NOVA_ERR_UNSUPPORTED, "FFT length must be 2^a 3^b: " + std::to_string(n));
    std::vector<double2> tw(n);
    for (int e = 0; e < n; e++) {
        long double ang = -2.0L * 3.141592653589793238462643383279502884L * (long double)e / (long double)n;
        tw[e] = make_double2((double)cosl(ang), (double)sinl(ang));
    }
    // digit-reversed load positions: walk the stages from the last one
    std::vector<int> pos(1, 0);
    std::vector<int> Ls(f.nr + 1, 1);
    for (int s = 0; s < f.nr; s++) Ls[s + 1] = Ls[s] * f.radix[s];
    for (int s = f.nr; s >= 1; s--) {
        const int cnt_prev = n / Ls[s - 1];
        const int cnt = n / Ls[s];
        std::vector<int> np(cnt_prev);
        for (int r = 0; r < cnt_prev; r++) np[r] = pos[r % cnt] + (r / cnt) * Ls[s - 1];
        pos.swap(np);
    }
    NOVA_HIP(hipMalloc((void **)&f.tw, sizeof(double2) * n));
    NOVA_HIP(hipMalloc((void **)&f.perm, sizeof(int) * n));
    NOVA_HIP(hipMemcpy(f.tw, tw.data(), sizeof(double2) * n, hipMemcpyHostToDevice));
    NOVA_HIP(hipMemcpy(f.perm, pos.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    return NOVA_OK;
}

void free_len(FftLen &f)
{
    if (f.tw) (void)hipFree(f.tw);
    if (f.perm) (void)hipFree(f.perm);
    f.tw = nullptr;
    f.perm = nullptr;
}

FftDev dev(const FftLen &f)
{
    FftDev d;
    d.n = f.n;
    d.nr = f.nr;
    for (int s = 0; s < 16; s++) d.radix[s] = f.radix[s];
    return d;
}

size_t smem_bytes(int n, int lines) { return sizeof(double2) * ((size_t)n + (n + 1) / 2 + (size_t)n * lines); }

constexpr size_t kLdsLimit = 64 * 1024;  // dynamic LDS per block without opt-in

int ylines_per_block(int n)
{
    int lb = 2048 / n;
    return lb < 1 ? 1 : (lb > 16 ? 16 : lb);
}

int xbatch(int n)
{
    int jb = 2048 / n;
    return jb < 1 ? 1 : (jb > 8 ? 8 : jb);
}

// poisson::wavenumbersFFT_ modified wavenumber (poisson.cpp:365-389).
double kmod(double kk, double h, const double *coef, int hm)
{
    double kp = 0.0;
    if (kk > kEps)
        for (int ll = 0; ll < hm; ll++) kp += kk * coef[ll] / (ll + 1) * std::sin((ll + 1) * kk * h) / (kk * h);
    return kp;
}

}  // namespace

int poisson_create(nova_ctx *ctx, nova_poisson **out)
{
    const Grid &g = ctx->g;
    if (g.wall[0] || g.wall[2]) return fail(NOVA_ERR_UNSUPPORTED, "Poisson walls are supported in y only");
    if (g.n[2] % 2) return fail(NOVA_ERR_UNSUPPORTED, "N_z must be even");
    nova_poisson *ps = new nova_poisson();
    ps->ctx = ctx;
    ps->Nx = g.n[0];
    ps->Ny = g.n[1];
    ps->Nz = g.n[2];
    ps->wall_y = g.wall[1];
    ps->sy = g.s[1];
    ps->Ty = ps->wall_y ? 2 * ps->Ny : ps->Ny;
    ps->Kc = ps->Nz / 2 + 1;
    ps->norm = (double)((long)ps->Nx * (long)ps->Ty * (long)ps->Nz);
    int rc;
    if ((rc = make_len(ps->fx, ps->Nx)) || (rc = make_len(ps->fy, ps->Ty)) || (rc = make_len(ps->fz, ps->Nz))) {
        free_len(ps->fx); free_len(ps->fy); free_len(ps->fz);
        delete ps;
        return rc;
    }
    if (smem_bytes(ps->Nz, kZLines / 2) > kLdsLimit || smem_bytes(ps->Nx, xbatch(ps->Nx)) > kLdsLimit ||
        smem_bytes(ps->Ty, ylines_per_block(ps->Ty)) > kLdsLimit) {
        free_len(ps->fx); free_len(ps->fy); free_len(ps->fz);
        delete ps;
        return fail(NOVA_ERR_UNSUPPORTED, "grid too large for the LDS line FFT");
    }
    const size_t nh = (size_t)ps->Kc * ps->Nx * ps->sy;
    NOVA_HIP(hipMalloc((void **)&ps->H, sizeof(double2) * nh));
    // modified wavenumbers per index, with the reference's wrap
    const double pi = 3.14159265358979323846;
    std::vector<double> kx(ps->Nx), ky(ps->sy), kz(ps->Kc);
    const double Ly = ps->wall_y ? 2.0 * ctx->desc.L[1] : ctx->desc.L[1];
    for (int i = 0; i < ps->Nx; i++) {
        double k = (2 * i < ps->Nx) ? 2.0 * pi * i / ctx->desc.L[0] : 2.0 * pi * (ps->Nx - i) / ctx->desc.L[0];
        kx[i] = kmod(k, ctx->d[0], ctx->coef, g.h);
    }
    for (int j = 0; j < ps->sy; j++) {
        double k = (2 * j < ps->Ty) ? 2.0 * pi * j / Ly : 2.0 * pi * (ps->Ty - j) / Ly;
        ky[j] = kmod(k, ctx->d[1], ctx->coef, g.h);
    }
    for (int k = 0; k < ps->Kc; k++) kz[k] = kmod(2.0 * pi * k / ctx->desc.L[2], ctx->d[2], ctx->coef, g.h);
    NOVA_HIP(hipMalloc((void **)&ps->kpx, sizeof(double) * (ps->Nx + ps->sy + ps->Kc)));
    ps->kpy = ps->kpx + ps->Nx;
    ps->kpz = ps->kpy + ps->sy;
    NOVA_HIP(hipMemcpy(ps->kpx, kx.data(), sizeof(double) * ps->Nx, hipMemcpyHostToDevice));
    NOVA_HIP(hipMemcpy(ps->kpy, ky.data(), sizeof(double) * ps->sy, hipMemcpyHostToDevice));
    NOVA_HIP(hipMemcpy(ps->kpz, kz.data(), sizeof(double) * ps->Kc, hipMemcpyHostToDevice));
    *out = ps;
    return NOVA_OK;
}

int poisson_destroy(nova_poisson *ps)
{
    if (!ps) return NOVA_OK;
    free_len(ps->fx);
    free_len(ps->fy);
    free_len(ps->fz);
    if (ps->H) (void)hipFree(ps->H);
    if (ps->kpx) (void)hipFree(ps->kpx);
    if (ps->ptmp) (void)hipFree(ps->ptmp);
    delete ps;
    return NOVA_OK;
}

// rhs: scalar (src_is_vector = 0) or the vector whose divergence*scale is
// the right-hand side.
int poisson_run(nova_poisson *ps, const double *src, int src_is_vector, double scale, double *p)
{
    nova_ctx *ctx = ps->ctx;
    const Grid &g = ctx->g;
    hipStream_t st = ctx->stream;
    const FftDev dz = dev(ps->fz), dy = dev(ps->fy), dx = dev(ps->fx);
    {
        dim3 grid((ps->sy + kZLines - 1) / kZLines, ps->Nx);
        TimedScope ts(ctx, src_is_vector ? "fft_div_r2c_z" : "fft_r2c_z");
        hipLaunchKernelGGL(k_fwd_z, grid, dim3(kFftThreads), smem_bytes(ps->Nz, kZLines / 2), st, g, dz, ps->fz.tw,
                           ps->fz.perm, src, src_is_vector, scale, ps->H, ps->Nx, ps->sy);
        if (int rc = check_launch("poisson fwd_z")) return rc;
    }
    const int64_t ylines = (int64_t)ps->Kc * ps->Nx;
    const int lb = ylines_per_block(ps->Ty);
    const unsigned yblocks = (unsigned)((ylines + lb - 1) / lb);
    {
    TimedScope ts(ctx, "fft_y");
    hipLaunchKernelGGL(k_fft_y, dim3(yblocks), dim3(kFftThreads), smem_bytes(ps->Ty, lb), st, dy, ps->fy.tw,
                       ps->fy.perm, ps->H, ylines, ps->sy, lb, ps->wall_y, 0);
    }
    if (int rc = check_launch("poisson fft_y")) return rc;
    const int jb = xbatch(ps->Nx);
    {
        dim3 grid((ps->sy + jb - 1) / jb, ps->Kc);
        TimedScope ts(ctx, "fft_x_solve");
        hipLaunchKernelGGL(k_fft_x, grid, dim3(kFftThreads), smem_bytes(ps->Nx, jb), st, dx, ps->fx.tw,
                           ps->fx.perm, ps->H, ps->Nx, ps->sy, jb, ps->kpx, ps->kpy, ps->kpz);
        if (int rc = check_launch("poisson fft_x")) return rc;
    }
    {
    TimedScope ts(ctx, "fft_y");
    hipLaunchKernelGGL(k_fft_y, dim3(yblocks), dim3(kFftThreads), smem_bytes(ps->Ty, lb), st, dy, ps->fy.tw,
                       ps->fy.perm, ps->H, ylines, ps->sy, lb, ps->wall_y, 1);
    }
    if (int rc = check_launch("poisson ifft_y")) return rc;
    {
        dim3 grid((ps->sy + kZLines - 1) / kZLines, ps->Nx);
        TimedScope ts(ctx, "fft_c2r_z");
        hipLaunchKernelGGL(k_inv_z, grid, dim3(kFftThreads), smem_bytes(ps->Nz, kZLines / 2), st, g, dz, ps->fz.tw,
                           ps->fz.perm, ps->H, p, ps->Nx, ps->sy, 1.0 / ps->norm);
        if (int rc = check_launch("poisson inv_z")) return rc;
    }
    return NOVA_OK;
}

int poisson_solve(nova_poisson *ps, const double *rhs, double *p)
{
    return poisson_run(ps, rhs, 0, 1.0, p);
}

int project(nova_poisson *ps, double *F, double s, double *p)
{
    nova_ctx *ctx = ps->ctx;
    if (!p) {
        if (!ps->ptmp) NOVA_HIP(hipMalloc((void **)&ps->ptmp, sizeof(double) * ctx->g.np));
        p = ps->ptmp;
    }
    if (int rc = poisson_run(ps, F, 1, 1.0 / s, p)) return rc;
    const Grid &g = ctx->g;
    dim3 grid((g.s[2] + 63) / 64, (g.s[1] + 3) / 4, g.s[0]);
    TimedScope ts(ctx, "grad_correct");
    hipLaunchKernelGGL(k_correct, grid, dim3(64, 4, 1), 0, ctx->stream, g, p, s, F);
    return check_launch("project correct");
}

}  // namespace nova

namespace nova {
nova_ctx *poisson_ctx(nova_poisson *ps) { return ps->ctx; }
}  // namespace nova

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

}  // namespace nova

struct nova_poisson {
    nova_ctx *ctx = nullptr;
    int Nx = 0, Ny = 0, Nz = 0;
    int sy = 0;      // stored y extent
    int wall_y = 0;  // 1: a walled y solved on the even extension
    int Ty = 0;      // y transform length
    int Kc = 0;      // Nz/2 + 1
    nova::FftLen fx, fy, fz;
    double2 *H = nullptr;
    double *kpx = nullptr, *kpy = nullptr, *kpz = nullptr;
    double norm = 1.0;
    double *ptmp = nullptr;  // potential when the caller passes none
    double *dtmp = nullptr;  // divergence scratch of the split projection
    int batch_cap = 1;       // fields one batched solve can take
    int scheme = NOVA_SCHEME_EXPLICIT;  // derivative scheme of div, grad and k'
};

namespace nova {

namespace {

// threads per block: the k passes (r2c, c2r) and the j and i passes (512
// threads made the j passes 4 us and the i pass 8 us faster at 256^3, and
// the c2r pass 3 us slower)
constexpr int kFftThreads = 256;
constexpr int kLineThreads = 512;
constexpr double kEps = 1.0e-16;  // src/Tools/Tools.cpp:7

__device__ inline double2 cmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ inline double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
// multiply by -i (forward) or +i (inverse)
__device__ inline double2 rot(double2 a, bool inv) { return inv ? make_double2(-a.y, a.x) : make_double2(a.y, -a.x); }

// ---- compile-time mixed-radix plans: radix 4 first, then 2, then 3 (the
// host-side digit-reversal table uses the same order) ----
constexpr int plan_nr(int n)
{
    int c = 0;
    while (n % 4 == 0) { n /= 4; c++; }
    while (n % 2 == 0) { n /= 2; c++; }
    while (n % 3 == 0) { n /= 3; c++; }
    return n == 1 ? c : -1;
}

constexpr int plan_radix(int n, int s)
{
    int idx = 0;
    while (n % 4 == 0) { if (idx == s) return 4; n /= 4; idx++; }
    while (n % 2 == 0) { if (idx == s) return 2; n /= 2; idx++; }
    while (n % 3 == 0) { if (idx == s) return 3; n /= 3; idx++; }
    return 1;
}

constexpr int plan_lp(int n, int s)
{
    int acc = 1;
    for (int t = 0; t < s; t++) acc *= plan_radix(n, t);
    return acc;
}

template <int R, bool INV>
__device__ inline void dft_small(double2 *a)
{
    if constexpr (R == 2) {
        const double2 t = a[0];
        a[0] = cadd(t, a[1]);
        a[1] = csub(t, a[1]);
    } else if constexpr (R == 4) {
        const double2 s02 = cadd(a[0], a[2]), d02 = csub(a[0], a[2]);
        const double2 s13 = cadd(a[1], a[3]), d13 = rot(csub(a[1], a[3]), INV);
        a[0] = cadd(s02, s13);
        a[2] = csub(s02, s13);
        a[1] = cadd(d02, d13);
        a[3] = csub(d02, d13);
    } else if constexpr (R == 3) {
        const double c = -0.5, s = 0.86602540378443864676;
        const double2 sum = cadd(a[1], a[2]);
        const double2 dif = csub(a[1], a[2]);
        const double2 t = make_double2(a[0].x + c * sum.x, a[0].y + c * sum.y);
        const double2 r = rot(make_double2(s * dif.x, s * dif.y), INV);  // -i s (a1 - a2) forward
        a[0] = cadd(a[0], sum);
        a[1] = cadd(t, r);
        a[2] = csub(t, r);
    }
}

// LDS position of element e of a line held contiguously (ES = 1): the low
// four bits are XORed with bits 2-5 ^ 6-9, a permutation inside each
// aligned 16-element block, so that the radix-4 passes (legs 1, 4, 16, 64
// elements apart), the digit-reversed load and the natural-order and
// mirrored reads each hit 16 distinct 16-byte bank groups in 16
// consecutive lanes, or nearly so.  Lengths that are not a multiple of 16
// keep the plain layout.
template <int N>
__device__ inline int swz(int e)
{
    if constexpr (N % 16 == 0) return e ^ (((e >> 2) ^ (e >> 6)) & 15);
    else return e;
}

// One radix pass over NLINES line FFTs of length N held in LDS; element e
// of line l at buf[l*LS + e*ES] (at buf[l*LS + swz(e)] when ES = 1).  DIT
// twiddles before the butterfly, DIF after it.  Consecutive threads walk
// the index that is contiguous in LDS.
template <int N, int S, bool DIT, bool INV, int NLINES, int LS, int ES, int NT>
__device__ inline void fft_pass(double2 *buf, const double2 *tw)
{
    constexpr int R = plan_radix(N, S);
    constexpr int Lp = plan_lp(N, S);
    constexpr int L = Lp * R;
    constexpr int NB = N / R;
    constexpr int TSTEP = N / L;
    constexpr int TOTAL = NB * NLINES;
    constexpr bool BATCH_FAST = ES != 1;
    for (int w = threadIdx.x; w < TOTAL; w += NT) {
        int line, bf;
        if constexpr (BATCH_FAST) { line = w % NLINES; bf = w / NLINES; }
        else { line = w / NB; bf = w % NB; }
        const int g = bf / Lp;
        const int q = bf % Lp;
        double2 *bl = buf + line * LS;
        const int e0 = g * L + q;
        auto at = [&](int p) -> double2 & {
            if constexpr (ES == 1) return bl[swz<N>(e0 + p * Lp)];
            else return bl[(e0 + p * Lp) * ES];
        };
        double2 a[R];
#pragma unroll
        for (int p = 0; p < R; p++) a[p] = at(p);
        // the pass with Lp = 1 has q = 0: every twiddle is tw[0] = (1, 0),
        // whose product leaves a value unchanged, so neither the table
        // reads nor the products are issued
        if constexpr (DIT) {
            if constexpr (Lp > 1) {
#pragma unroll
                for (int p = 1; p < R; p++) {
                    double2 wv = tw[p * q * TSTEP];
                    if (INV) wv.y = -wv.y;
                    a[p] = cmul(a[p], wv);
                }
            }
            dft_small<R, INV>(a);
        } else {
            dft_small<R, INV>(a);
            if constexpr (Lp > 1) {
#pragma unroll
                for (int p = 1; p < R; p++) {
                    double2 wv = tw[p * q * TSTEP];
                    if (INV) wv.y = -wv.y;
                    a[p] = cmul(a[p], wv);
                }
            }
        }
#pragma unroll
        for (int p = 0; p < R; p++) at(p) = a[p];
    }
    __syncthreads();
}

// All passes: DIT expects digit-reversed input and leaves natural order;
// DIF takes natural order and leaves digit-reversed output.
template <int N, bool DIT, bool INV, int NLINES, int LS, int ES, int NT, int IT = 0>
__device__ inline void lds_fft(double2 *buf, const double2 *tw)
{
    constexpr int NR = plan_nr(N);
    static_assert(NR > 0, "FFT length must be 2^a 3^b");
    if constexpr (IT < NR) {
        fft_pass<N, DIT ? IT : NR - 1 - IT, DIT, INV, NLINES, LS, ES, NT>(buf, tw);
        lds_fft<N, DIT, INV, NLINES, LS, ES, NT, IT + 1>(buf, tw);
    }
}

template <int N, int NT>
__device__ inline void load_tables(double2 *stw, int *sperm, const double2 *tw, const int *perm)
{
    for (int e = threadIdx.x; e < N; e += NT) {
        stw[e] = tw[e];
        sperm[e] = perm[e];
    }
    __syncthreads();
}

__device__ inline int64_t lin(const Grid &g, int i, int j, int k)
{
    return ((int64_t)i * g.s[1] + j) * g.s[2] + k;
}

// Stored index of node q along a direction: one wrap suffices on a periodic
// axis (|offset| <= 4 < n); walls reflect.
template <bool WALLS>
__device__ inline int wrapq(int q, int n, int wall, int &flip)
{
    if (!WALLS) {
        flip = 0;
        q += q < 0 ? n : 0;
        q -= q >= n ? n : 0;
        return q;
    }
    return bmap(q, n, wall, flip);
}

template <bool WALLS>
__device__ inline double ldw(const double *__restrict__ f, const Grid &g, int nc, int c, int i, int j, int k)
{
    int fi, fj, fk;
    const int ii = wrapq<WALLS>(i, g.n[0], g.wall[0], fi);
    const int jj = wrapq<WALLS>(j, g.n[1], g.wall[1], fj);
    const int kk = wrapq<WALLS>(k, g.n[2], g.wall[2], fk);
    const double v = f[c * g.np + lin(g, ii, jj, kk)];
    if (!WALLS) return v;
    return bsign(g, nc, c, 0, fi) * bsign(g, nc, c, 1, fj) * bsign(g, nc, c, 2, fk) * v;
}

// div(F) * scale at a stored node (src/ex/div/div.hxx:37-77).
template <bool WALLS>
__device__ inline double div_at(const Grid &g, const double *__restrict__ F, int i, int j, int k, double scale)
{
    double ax = 0.0, ay = 0.0, az = 0.0;
    for (int l = 0; l < g.h; l++) {
        ax += g.a[0][l] * (ldw<WALLS>(F, g, 3, 0, i + 1 + l, j, k) - ldw<WALLS>(F, g, 3, 0, i - 1 - l, j, k));
        ay += g.a[1][l] * (ldw<WALLS>(F, g, 3, 1, i, j + 1 + l, k) - ldw<WALLS>(F, g, 3, 1, i, j - 1 - l, k));
        az += g.a[2][l] * (ldw<WALLS>(F, g, 3, 2, i, j, k + 1 + l) - ldw<WALLS>(F, g, 3, 2, i, j, k - 1 - l));
    }
    return (ax + ay + az * g.zinv[k]) * scale;
}

// Real lines per block in the k passes.
template <int N>
constexpr int zlines() { return (N >= 128 && N <= 256) ? 8 : (N <= 384 ? 16 : (N <= 512 ? 8 : 4)); }
// The r2c pass at 256: 16 real lines on 512 threads (68.7 against 72.1 us;
// the c2r pass measured 10 us slower so, and keeps 8 lines on 256).
template <int N>
constexpr int zlines_fwd() { return N == 256 ? 16 : zlines<N>(); }
template <int N>
constexpr int zthreads_fwd() { return N == 256 ? 512 : kFftThreads; }
// Lines per block along j, and batch width along i.
template <int N>
constexpr int ylines() { return N >= 2048 ? 1 : (2048 / N > 16 ? 16 : 2048 / N); }
template <int N>
constexpr int xbatch_t() { return N >= 2048 ? 1 : (2048 / N > 8 ? 8 : 2048 / N); }

// pass 1: source -> half spectrum H[kc][i][j].
template <int N, bool WALLS>
__global__ void __launch_bounds__(zthreads_fwd<N>()) k_fwd_z(Grid g, const double2 *__restrict__ tw,
                                                       const int *__restrict__ perm, const double *__restrict__ src,
                                                       int src_is_vector, double scale, double2 *__restrict__ H,
                                                       int Nx, int sy, int64_t src_bs)
{
    constexpr int ZL = zlines_fwd<N>();
    constexpr int NT = zthreads_fwd<N>();
    // batch member blockIdx.z: its source and half spectrum
    src += blockIdx.z * src_bs;
    H += (int64_t)blockIdx.z * (N / 2 + 1) * Nx * sy;
    __shared__ double2 stw[N];
    __shared__ int sperm[N];
    __shared__ double2 buf[(ZL / 2) * N];
    load_tables<N, NT>(stw, sperm, tw, perm);
    const int i = blockIdx.y;
    const int j0 = blockIdx.x * ZL;
    if (src_is_vector) {
        for (int e = threadIdx.x; e < ZL * N; e += NT) {
            const int jj = e / N, k = e % N;
            const int j = j0 + jj;
            double *slot = (double *)&buf[(jj >> 1) * N + swz<N>(sperm[k])];
            slot[jj & 1] = j < sy ? div_at<WALLS>(g, src, i, j, k, scale) : 0.0;
        }
    } else {
        // A lane issues all its loads before its first LDS write: a loop
        // that writes each value as it lands keeps one load per wave in
        // flight, and the pass waits on memory latency.  The scalar source
        // is read once: a non-temporal load leaves the last-level cache to
        // the half spectrum written here, which the j pass reads next.
        // The loads are unconditional (rows past the grid read its last
        // row, zeroed below), so that no branch separates them.
        constexpr int PER = (ZL * N + NT - 1) / NT;
        double v[PER];
#pragma unroll
        for (int r = 0; r < PER; r++) {
            const int e = threadIdx.x + r * NT;
            const int j = j0 + e / N;
            v[r] = __builtin_nontemporal_load(&src[lin(g, i, j < sy ? j : sy - 1, e % N)]);
        }
#pragma unroll
        for (int r = 0; r < PER; r++) {
            const int e = threadIdx.x + r * NT;
            const int jj = e / N, k = e % N;
            if (e < ZL * N) ((double *)&buf[(jj >> 1) * N + swz<N>(sperm[k])])[jj & 1] = j0 + jj < sy ? v[r] * scale : 0.0;
        }
    }
    __syncthreads();
    lds_fft<N, true, false, ZL / 2, N, 1, NT>(buf, stw);
    constexpr int Kc = N / 2 + 1;
    for (int e = threadIdx.x; e < ZL * Kc; e += NT) {
        const int jj = e % ZL, kc = e / ZL;
        const int j = j0 + jj;
        if (j >= sy) continue;
        const double2 *Z = buf + (jj >> 1) * N;
        const double2 z = Z[swz<N>(kc)], zm = Z[swz<N>((N - kc) % N)];
        double2 out;
        if ((jj & 1) == 0) out = make_double2(0.5 * (z.x + zm.x), 0.5 * (z.y - zm.y));
        else out = make_double2(0.5 * (z.y + zm.y), -0.5 * (z.x - zm.x));
        H[((int64_t)kc * Nx + i) * sy + j] = out;
    }
}

// passes 2 and 4: lines along j (contiguous).  EXT = even extension of the
// sy = N_y + 1 stored nodes to length N = 2 N_y (DCT-I through the FFT).
template <int N, bool EXT, bool INV>
__global__ void __launch_bounds__(kLineThreads) k_fft_y(const double2 *__restrict__ tw, const int *__restrict__ perm,
                                                       double2 *__restrict__ H, int64_t nlines_total, int sy)
{
    constexpr int LB = ylines<N>();
    __shared__ double2 stw[N];
    __shared__ int sperm[N];
    __shared__ double2 buf[LB * N];
    load_tables<N, kLineThreads>(stw, sperm, tw, perm);
    const int64_t l0 = (int64_t)blockIdx.x * LB;
    constexpr int Ny = EXT ? N / 2 : N;
    // all of a lane's loads in flight before its LDS writes
    constexpr int PER = (LB * N + kLineThreads - 1) / kLineThreads;
    double2 v[PER];
#pragma unroll
    for (int r = 0; r < PER; r++) {
        const int e = threadIdx.x + r * kLineThreads;
        const int ln = e / N, q = e % N;
        const int64_t L = l0 + ln;
        v[r] = make_double2(0.0, 0.0);
        if (e < LB * N && L < nlines_total) {
            const int src = (EXT && q > Ny) ? 2 * Ny - q : q;
            v[r] = H[L * sy + src];
        }
    }
#pragma unroll
    for (int r = 0; r < PER; r++) {
        const int e = threadIdx.x + r * kLineThreads;
        if (e < LB * N) buf[(e / N) * N + swz<N>(sperm[e % N])] = v[r];
    }
    __syncthreads();
    lds_fft<N, true, INV, LB, N, 1, kLineThreads>(buf, stw);
    // the stored part of each line: sy entries of the extension, or the N of
    // a periodic line (a walled y solved periodically over its first N
    // nodes leaves entry N, the wall node's, untouched)
    const int qmax = sy < N ? sy : N;
    for (int e = threadIdx.x; e < LB * qmax; e += kLineThreads) {
        const int ln = e / qmax, q = e % qmax;
        const int64_t L = l0 + ln;
        if (L < nlines_total) H[L * sy + q] = buf[ln * N + swz<N>(q)];
    }
}

// pass 3: lines along i, a batch of consecutive j; forward, divide by
// -k'^2, inverse.
template <int N>
__global__ void __launch_bounds__(kLineThreads) k_fft_x(const double2 *__restrict__ tw, const int *__restrict__ perm,
                                                       double2 *__restrict__ H, int sy,
                                                       const double *__restrict__ kpx, const double *__restrict__ kpy,
                                                       const double *__restrict__ kpz, int Kc)
{
    constexpr int JB = xbatch_t<N>();
    __shared__ double2 stw[N];
    __shared__ int sperm[N];
    __shared__ double2 buf[N * JB];
    load_tables<N, kLineThreads>(stw, sperm, tw, perm);
    // blockIdx.y runs over (batch member, kc) of the [b][kc][i][j] layout
    const int kc = blockIdx.y % Kc;
    const int j0 = blockIdx.x * JB;
    double2 *Hs = H + (int64_t)blockIdx.y * N * sy;
    // all of a lane's loads in flight before its LDS writes
    constexpr int PER = (N * JB + kLineThreads - 1) / kLineThreads;
    double2 v[PER];
#pragma unroll
    for (int r = 0; r < PER; r++) {
        const int e = threadIdx.x + r * kLineThreads;
        const int jj = e % JB, i = e / JB;
        const int j = j0 + jj;
        v[r] = make_double2(0.0, 0.0);
        if (e < N * JB && j < sy) v[r] = Hs[(int64_t)i * sy + j];
    }
    const double kz = kpz[kc];
    // the wavenumbers of a lane's elements, loaded with the
    // spectrum, so that they land during the forward transform (clamped
    // indices past the end are never used)
    double kxr[PER], kyr[PER];
#pragma unroll
    for (int r = 0; r < PER; r++) {
        const int e = threadIdx.x + r * kLineThreads;
        const int j = j0 + e % JB;
        kxr[r] = kpx[e < N * JB ? e / JB : 0];
        kyr[r] = kpy[j < sy ? j : 0];
    }
#pragma unroll
    for (int r = 0; r < PER; r++) {
        const int e = threadIdx.x + r * kLineThreads;
        if (e < N * JB) buf[sperm[e / JB] * JB + e % JB] = v[r];
    }
    __syncthreads();
    lds_fft<N, true, false, JB, 1, JB, kLineThreads>(buf, stw);
#pragma unroll
    for (int r = 0; r < PER; r++) {
        const int e = threadIdx.x + r * kLineThreads;
        const int jj = e % JB, i = e / JB;
        const int j = j0 + jj;
        if (e >= N * JB || j >= sy) continue;
        double k2;
        {
#pragma clang fp contract(off)
            const double kx = kxr[r], ky = kyr[r];
            k2 = kx * kx + ky * ky + kz * kz;
        }
        double2 v = buf[i * JB + jj];
        if (k2 < kEps) v = make_double2(0.0, 0.0);
        else v = make_double2(-v.x / k2, -v.y / k2);
        buf[i * JB + jj] = v;
    }
    __syncthreads();
    lds_fft<N, false, true, JB, 1, JB, kLineThreads>(buf, stw);
    for (int e = threadIdx.x; e < N * JB; e += kLineThreads) {
        const int jj = e % JB, i = e / JB;
        const int j = j0 + jj;
        if (j < sy) Hs[(int64_t)i * sy + j] = buf[sperm[i] * JB + jj];
    }
}

// pass 5: half spectrum -> real lines along k, p = v / norm.
template <int N>
__global__ void __launch_bounds__(kFftThreads) k_inv_z(Grid g, const double2 *__restrict__ tw,
                                                       const int *__restrict__ perm, const double2 *__restrict__ H,
                                                       double *__restrict__ p, int Nx, int sy, double inv_norm,
                                                       int64_t p_bs)
{
    constexpr int ZL = zlines<N>();
    H += (int64_t)blockIdx.z * (N / 2 + 1) * Nx * sy;
    p += blockIdx.z * p_bs;
    __shared__ double2 stw[N];
    __shared__ int sperm[N];
    __shared__ double2 buf[(ZL / 2) * N];
    load_tables<N, kFftThreads>(stw, sperm, tw, perm);
    const int i = blockIdx.y;
    const int j0 = blockIdx.x * ZL;
    constexpr int Kc = N / 2 + 1;
    // Z = A + iB, Hermitian-extended: A from the even j, B from the odd j of
    // a pair; a lane reads both (32 contiguous bytes)
    // (all of a lane's loads in flight before its LDS writes)
    constexpr int PER = ((ZL / 2) * Kc + kFftThreads - 1) / kFftThreads;
    double2 v0[PER], v1[PER];
#pragma unroll
    for (int r = 0; r < PER; r++) {
        const int e = threadIdx.x + r * kFftThreads;
        const int pr = e % (ZL / 2), kc = e / (ZL / 2);
        const int j = j0 + 2 * pr;
        const int64_t row = ((int64_t)kc * Nx + i) * sy;
        v0[r] = v1[r] = make_double2(0.0, 0.0);
        if (e < (ZL / 2) * Kc && j < sy) v0[r] = H[row + j];
        if (e < (ZL / 2) * Kc && j + 1 < sy) v1[r] = H[row + j + 1];
    }
#pragma unroll
    for (int r = 0; r < PER; r++) {
        const int e = threadIdx.x + r * kFftThreads;
        const int pr = e % (ZL / 2), kc = e / (ZL / 2);
        if (e >= (ZL / 2) * Kc) continue;
        double2 *Z = buf + pr * N;
        const bool mirror = kc > 0 && 2 * kc < N;
        Z[swz<N>(sperm[kc])] = make_double2(v0[r].x - v1[r].y, v0[r].y + v1[r].x);
        if (mirror) Z[swz<N>(sperm[N - kc])] = make_double2(v0[r].x + v1[r].y, -v0[r].y + v1[r].x);  // + i conj(v1)
    }
    __syncthreads();
    lds_fft<N, true, true, ZL / 2, N, 1, kFftThreads>(buf, stw);
    for (int e = threadIdx.x; e < ZL * N; e += kFftThreads) {
        const int jj = e / N, k = e % N;
        const int j = j0 + jj;
        if (j >= sy) continue;
        const double2 z = buf[(jj >> 1) * N + swz<N>(k)];
        p[lin(g, i, j, k)] = ((jj & 1) ? z.y : z.x) * inv_norm;
    }
}

// F <- F - s * grad(p) (app/TG_Piosson/RKLoop.H:36, app/OT_Poisson/BEqn.H:40).
template <bool WALLS>
__global__ void k_correct(Grid g, const double *__restrict__ p, double s, double *__restrict__ F)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int i = blockIdx.z;
    if (k >= g.s[2] || j >= g.s[1]) return;
    double gx = 0.0, gy = 0.0, gz = 0.0;
    for (int l = 0; l < g.h; l++) {
        gx += g.a[0][l] * (ldw<WALLS>(p, g, 1, 0, i + 1 + l, j, k) - ldw<WALLS>(p, g, 1, 0, i - 1 - l, j, k));
        gy += g.a[1][l] * (ldw<WALLS>(p, g, 1, 0, i, j + 1 + l, k) - ldw<WALLS>(p, g, 1, 0, i, j - 1 - l, k));
        gz += g.a[2][l] * (ldw<WALLS>(p, g, 1, 0, i, j, k + 1 + l) - ldw<WALLS>(p, g, 1, 0, i, j, k - 1 - l));
    }
    gz *= g.zinv[k];
    const int64_t q = lin(g, i, j, k);
    F[q] = F[q] - s * gx;
    F[g.np + q] = F[g.np + q] - s * gy;
    F[2 * g.np + q] = F[2 * g.np + q] - s * gz;
}

// ---- streamed divergence and gradient correction ----
//
// One thread per (j,k) column walks i over a chunk of planes; the i
// neighbours of its column stay in a register queue, the j and k neighbours
// come from global memory (shared L1/L2 lines between adjacent rows).
constexpr int kRK = 64, kRJ = 4;

struct RowTiles {
    int ichunk, nkt, njt, nic;
};

// A batch of same-grid fields: member b at F + b f_bs (potential p + b p_bs);
// members b >= split use the second scale.
struct BatchArgs {
    int64_t f_bs = 0, p_bs = 0;
    int split = 1;
    double scale_hi = 1.0, s_hi = 1.0;
};

__device__ inline bool row_tile(const Grid &g, const RowTiles &t, int &k, int &j, int &i0, int &i1)
{
    const int T = t.nkt * t.njt * t.nic;
    int b = blockIdx.x;
    if (T % 8 == 0) b = (blockIdx.x % 8) * (T / 8) + blockIdx.x / 8;
    const int kt = b % t.nkt;
    const int jt = (b / t.nkt) % t.njt;
    const int ic = b / (t.nkt * t.njt);
    k = kt * kRK + threadIdx.x;
    j = jt * kRJ + threadIdx.y;
    i0 = ic * t.ichunk;
    i1 = i0 + t.ichunk;
    if (i1 > g.s[0]) i1 = g.s[0];
    return k < g.s[2] && j < g.s[1];
}

// Loop-invariant neighbour offsets of a (j,k) column: column offsets of
// the j+-(l+1) and k+-(l+1) neighbours and their reflection flags.
template <int H, bool WALLS>
struct ColNbrs {
    int jp[H], jm[H], kp[H], km[H];
    unsigned fl;  // bit 4l+0..3: flips of jp, jm, kp, km
    __device__ inline void init(const Grid &g, int j, int k)
    {
        fl = 0;
        const int s2 = g.s[2];
#pragma unroll
        for (int l = 0; l < H; l++) {
            int f0, f1, f2, f3;
            jp[l] = wrapq<WALLS>(j + 1 + l, g.n[1], g.wall[1], f0) * s2 + k;
            jm[l] = wrapq<WALLS>(j - 1 - l, g.n[1], g.wall[1], f1) * s2 + k;
            kp[l] = j * s2 + wrapq<WALLS>(k + 1 + l, g.n[2], g.wall[2], f2);
            km[l] = j * s2 + wrapq<WALLS>(k - 1 - l, g.n[2], g.wall[2], f3);
            fl |= (unsigned)(f0 | (f1 << 1) | (f2 << 2) | (f3 << 3)) << (4 * l);
        }
    }
    __device__ inline double sg(int l, int which) const
    {
        return (WALLS && ((fl >> (4 * l + which)) & 1u)) ? -1.0 : 1.0;
    }
};

// out = div(F) * scale (src/ex/div/div.hxx:37-77).  All loads of plane i+1
// are issued before plane i is combined.
template <int H, bool WALLS>
__global__ void __launch_bounds__(kRK *kRJ) k_div_rows(Grid g, RowTiles t, const double *__restrict__ F,
                                                       double scale, double *__restrict__ out, BatchArgs ba)
{
    int k, j, i0, i1;
    if (!row_tile(g, t, k, j, i0, i1)) return;
    // batch member blockIdx.y: field, output and scale group
    F += blockIdx.y * ba.f_bs;
    out += blockIdx.y * g.np;
    if ((int)blockIdx.y >= ba.split) scale = ba.scale_hi;
    const int64_t plane = (int64_t)g.s[1] * g.s[2];
    const int col = j * g.s[2] + k;
    ColNbrs<H, WALLS> nb;
    nb.init(g, j, k);
    const double *Fx = F, *Fy = F + g.np, *Fz = F + 2 * g.np;
    // F_y and F_z change sign through a reflection when they are odd there
    const bool oy = (g.odd[1] >> 1) & 1u, oz = g.wall[2] && ((g.odd[2] >> 2) & 1u);
    double q[2 * H + 1];
#pragma unroll
    for (int o = 0; o < 2 * H + 1; o++) {
        int fi;
        const int im = wrapq<WALLS>(i0 - H + o, g.n[0], g.wall[0], fi);
        q[o] = (WALLS && fi && (g.odd[0] & 1u) ? -1.0 : 1.0) * Fx[(int64_t)im * plane + col];
    }
    double vy[2 * H], wz[2 * H];
    auto load_plane = [&](int i, double (&y)[2 * H], double (&z)[2 * H]) {
        const int64_t pi = (int64_t)i * plane;
#pragma unroll
        for (int l = 0; l < H; l++) {
            y[l] = (oy ? nb.sg(l, 0) : 1.0) * Fy[pi + nb.jp[l]];
            y[H + l] = (oy ? nb.sg(l, 1) : 1.0) * Fy[pi + nb.jm[l]];
            z[l] = (oz ? nb.sg(l, 2) : 1.0) * Fz[pi + nb.kp[l]];
            z[H + l] = (oz ? nb.sg(l, 3) : 1.0) * Fz[pi + nb.km[l]];
        }
    };
    load_plane(i0, vy, wz);
    const double zi = g.zinv[k];
    for (int i = i0; i < i1; i++) {
        double nq = 0.0, ny[2 * H], nz[2 * H];
        if (i + 1 < i1) {
            int fi;
            const int im = wrapq<WALLS>(i + H + 1, g.n[0], g.wall[0], fi);
            nq = (WALLS && fi && (g.odd[0] & 1u) ? -1.0 : 1.0) * Fx[(int64_t)im * plane + col];
            load_plane(i + 1, ny, nz);
        }
        double ax = 0.0, ay = 0.0, az = 0.0;
#pragma unroll
        for (int l = 0; l < H; l++) {
            ax += g.a[0][l] * (q[H + 1 + l] - q[H - 1 - l]);
            ay += g.a[1][l] * (vy[l] - vy[H + l]);
            az += g.a[2][l] * (wz[l] - wz[H + l]);
        }
        out[(int64_t)i * plane + col] = (ax + ay + az * zi) * scale;
        if (i + 1 < i1) {
#pragma unroll
            for (int o = 0; o < 2 * H; o++) q[o] = q[o + 1];
            q[2 * H] = nq;
#pragma unroll
            for (int o = 0; o < 2 * H; o++) {
                vy[o] = ny[o];
                wz[o] = nz[o];
            }
        }
    }
}

// F <- F - s * grad(p) (app/TG_Piosson/RKLoop.H:36, app/OT_Poisson/BEqn.H:40).
template <int H, bool WALLS>
__global__ void __launch_bounds__(kRK *kRJ) k_correct_rows(Grid g, RowTiles t, const double *__restrict__ p,
                                                           double s, double *__restrict__ F, BatchArgs ba)
{
    int k, j, i0, i1;
    if (!row_tile(g, t, k, j, i0, i1)) return;
    F += blockIdx.y * ba.f_bs;
    p += blockIdx.y * ba.p_bs;
    if ((int)blockIdx.y >= ba.split) s = ba.s_hi;
    const int64_t plane = (int64_t)g.s[1] * g.s[2];
    const int col = j * g.s[2] + k;
    ColNbrs<H, WALLS> nb;
    nb.init(g, j, k);
    double q[2 * H + 1];
#pragma unroll
    for (int o = 0; o < 2 * H + 1; o++) {
        int fi;
        const int im = wrapq<WALLS>(i0 - H + o, g.n[0], g.wall[0], fi);
        q[o] = p[(int64_t)im * plane + col];
    }
    double py[2 * H], pz[2 * H], f[3];
    auto load_plane = [&](int i, double (&y)[2 * H], double (&z)[2 * H], double (&ff)[3]) {
        const int64_t pi = (int64_t)i * plane;
#pragma unroll
        for (int l = 0; l < H; l++) {
            y[l] = p[pi + nb.jp[l]];
            y[H + l] = p[pi + nb.jm[l]];
            z[l] = p[pi + nb.kp[l]];
            z[H + l] = p[pi + nb.km[l]];
        }
#pragma unroll
        for (int c = 0; c < 3; c++) ff[c] = F[c * g.np + pi + col];
    };
    load_plane(i0, py, pz, f);
    const double zi = g.zinv[k];
    for (int i = i0; i < i1; i++) {
        double nq = 0.0, ny[2 * H], nz[2 * H], nf[3];
        if (i + 1 < i1) {
            int fi;
            const int im = wrapq<WALLS>(i + H + 1, g.n[0], g.wall[0], fi);
            nq = p[(int64_t)im * plane + col];
            load_plane(i + 1, ny, nz, nf);
        }
        double gx = 0.0, gy = 0.0, gz = 0.0;
#pragma unroll
        for (int l = 0; l < H; l++) {
            gx += g.a[0][l] * (q[H + 1 + l] - q[H - 1 - l]);
            gy += g.a[1][l] * (py[l] - py[H + l]);
            gz += g.a[2][l] * (pz[l] - pz[H + l]);
        }
        gz *= zi;
        const int64_t q0 = (int64_t)i * plane + col;
        F[q0] = f[0] - s * gx;
        F[g.np + q0] = f[1] - s * gy;
        F[2 * g.np + q0] = f[2] - s * gz;
        if (i + 1 < i1) {
#pragma unroll
            for (int o = 0; o < 2 * H; o++) q[o] = q[o + 1];
            q[2 * H] = nq;
#pragma unroll
            for (int o = 0; o < 2 * H; o++) {
                py[o] = ny[o];
                pz[o] = nz[o];
            }
#pragma unroll
            for (int c = 0; c < 3; c++) f[c] = nf[c];
        }
    }
}

// ---- LDS-staged streamed divergence and gradient correction ----
//
// A 32 x 16 (k x j) tile walks i over a chunk like the stage kernels: the
// i-neighbours of its column sit in a register queue, the j and k
// neighbours come from LDS rows and columns staged per plane, and the loads
// of the next two planes are in flight while a plane is combined.  Values
// are kept as loaded and the wall reflection signs applied when they are
// staged.  DIV: out = div(F) * scale (the arithmetic of k_div_rows);
// otherwise F -= s grad(p) (that of k_correct_rows).
// (512 threads: against 32 x 8 the loads per node of a staged plane drop
// from 2.08 to 1.63 for p with both halos and from 1.47 to 1.28 for F_y,
// F_z with one halo each; measured 4% faster divergence and 3% faster
// correction at 256^3)
// Results are stored non-temporally (written once, read by the next kernel
// from HBM): at 256^3 the correction, the divergence and the stage that
// follows each ran 0.5-10% faster than with ordinary stores.
constexpr int kPK = 32, kPJ = 16;

template <int H, bool DIV, bool WALLS>
__global__ void __launch_bounds__(kPK *kPJ) k_proj_stream(Grid g, RowTiles t, const double *__restrict__ src, double sc,
                                                     double *__restrict__ dst, BatchArgs ba)
{
    constexpr int NH = 2 * H * kPK + kPJ * 2 * H;  // halo elements per plane
    static_assert(NH <= kPK * kPJ, "one halo slot per thread");
    const int T = t.nkt * t.njt * t.nic;
    int bt = blockIdx.x;
    if (T % 8 == 0) bt = (blockIdx.x % 8) * (T / 8) + blockIdx.x / 8;
    const int kt = bt % t.nkt, jt = (bt / t.nkt) % t.njt, ic = bt / (t.nkt * t.njt);
    const int tk = threadIdx.x % kPK, tj = threadIdx.x / kPK;
    const int k = kt * kPK + tk, j = jt * kPJ + tj;
    const bool valid = k < g.s[2] && j < g.s[1];
    const int i0 = ic * t.ichunk;
    const int i1 = i0 + t.ichunk < g.s[0] ? i0 + t.ichunk : g.s[0];
    const int64_t plane = (int64_t)g.s[1] * g.s[2];
    // batch member blockIdx.y
    if (DIV) {
        src += blockIdx.y * ba.f_bs;
        dst += blockIdx.y * g.np;
        if ((int)blockIdx.y >= ba.split) sc = ba.scale_hi;
    } else {
        src += blockIdx.y * ba.p_bs;
        dst += blockIdx.y * ba.f_bs;
        if ((int)blockIdx.y >= ba.split) sc = ba.s_hi;
    }
    const double *Fx = src, *Fy = src + g.np, *Fz = src + 2 * g.np;  // DIV only

    __shared__ double sa[kPJ + 2 * H][kPK + 2 * H];  // DIV: F_y (j halo); else p (both halos)
    __shared__ double sb[DIV ? kPJ : 1][DIV ? kPK + 2 * H : 1];  // DIV: F_z (k halo)

    int fj, fk;
    const int jm = bmap(j, g.n[1], g.wall[1], fj);
    const int km = bmap(k, g.n[2], g.wall[2], fk);
    const int64_t col = (int64_t)jm * g.s[2] + km;
    // this thread's halo slot: kind 1 = j row (F_y or p), 2 = k column (F_z or p)
    int hk = 0, hr = 0, hc = 0, hf = 0;
    int64_t hsrc = 0;
    {
        const int e = threadIdx.x;
        int jj = 0, kk = 0;
        if (e < 2 * H * kPK) {
            const int rr = e / kPK, cc = e % kPK;
            hk = 1;
            hr = rr < H ? rr : kPJ + rr;
            hc = cc + H;
            jj = jt * kPJ + (rr < H ? rr - H : kPJ + rr - H);
            kk = kt * kPK + cc;
        } else if (e < NH) {
            const int e2 = e - 2 * H * kPK;
            const int rr = e2 / (2 * H), c2 = e2 % (2 * H);
            hk = 2;
            hr = rr + H;
            hc = c2 < H ? c2 : kPK + c2;
            jj = jt * kPJ + rr;
            kk = kt * kPK + (c2 < H ? c2 - H : kPK + c2 - H);
        }
        int f1 = 0, f2 = 0;
        const int jm2 = bmap(jj, g.n[1], g.wall[1], f1);
        const int km2 = bmap(kk, g.n[2], g.wall[2], f2);
        hsrc = (int64_t)jm2 * g.s[2] + km2;
        hf = hk == 1 ? f1 : f2;  // the reflection that matters for the staged component
    }
    // sign of a staged value: DIV stages F_y (reflected across y walls) and
    // F_z (across z walls), each odd there or not as Grid::odd says; p is even
    const bool ox = g.odd[0] & 1u, oy = (g.odd[1] >> 1) & 1u, oz = (g.odd[2] >> 2) & 1u;
    auto sgn_y = [&](int flip) -> double { return (DIV && WALLS && flip && oy) ? -1.0 : 1.0; };
    auto sgn_z = [&](int flip) -> double { return (DIV && WALLS && flip && oz) ? -1.0 : 1.0; };

    struct Pf {
        double q;      // queue entry (F_x or p) of plane p + H
        double cy, cz; // DIV: F_y, F_z at the own column of plane p
        double h;      // halo value of plane p
        double f[3];   // correction: F at the own column of plane p
        int fq, fp;    // reflection flags of the queue plane and of plane p
    };
    auto plane_of = [&](int ii, int &flip) -> int64_t {
        const int im = bmap(ii, g.n[0], g.wall[0], flip);
        return (int64_t)im * plane;
    };
    const double *qsrc = DIV ? Fx : src;
    auto load_pf = [&](Pf &P, int p, bool queue) {
        if (queue) {
            const int64_t po = plane_of(p + H, P.fq);
            P.q = qsrc[po + col];
        }
        const int64_t pp = plane_of(p, P.fp);
        if (DIV) {
            P.cy = Fy[pp + col];
            P.cz = Fz[pp + col];
            P.h = hk == 1 ? Fy[pp + hsrc] : (hk == 2 ? Fz[pp + hsrc] : 0.0);
        } else {
            P.h = hk ? src[pp + hsrc] : 0.0;
            const int64_t qq = (int64_t)p * plane + col;
            if (valid)
#pragma unroll
                for (int c = 0; c < 3; c++) P.f[c] = dst[c * g.np + qq];
        }
    };

    double q[2 * H + 1];
#pragma unroll
    for (int o = 0; o < 2 * H + 1; o++) {
        int fi;
        const int64_t po = plane_of(i0 - H + o, fi);
        q[o] = (DIV && WALLS && fi && ox ? -1.0 : 1.0) * qsrc[po + col];
    }
    const double zi = g.zinv[km];
    const double cyz_sign_y = sgn_y(fj), cyz_sign_z = sgn_z(fk);

    auto body = [&](int i, Pf &Cur, Pf &Nxt) {
        // stage plane i
        if (DIV) {
            sa[tj + H][tk + H] = cyz_sign_y * Cur.cy;
            sb[tj][tk + H] = cyz_sign_z * Cur.cz;
            if (hk == 1) sa[hr][hc] = sgn_y(hf) * Cur.h;
            else if (hk == 2) sb[hr - H][hc] = sgn_z(hf) * Cur.h;
        } else {
            sa[tj + H][tk + H] = q[H];
            if (hk) sa[hr][hc] = Cur.h;
        }
        __syncthreads();
        const double f0 = Cur.f[0], f1 = Cur.f[1], f2 = Cur.f[2];
        if (i + 2 < i1) load_pf(Cur, i + 2, true);
        if (valid) {
            const int64_t q0 = (int64_t)i * plane + col;
            if (DIV) {
                double ax = 0.0, ay = 0.0, az = 0.0;
#pragma unroll
                for (int l = 0; l < H; l++) {
                    ax += g.a[0][l] * (q[H + 1 + l] - q[H - 1 - l]);
                    ay += g.a[1][l] * (sa[tj + H + 1 + l][tk + H] - sa[tj + H - 1 - l][tk + H]);
                    az += g.a[2][l] * (sb[tj][tk + H + 1 + l] - sb[tj][tk + H - 1 - l]);
                }
                __builtin_nontemporal_store((ax + ay + az * zi) * sc, dst + q0);
            } else {
                double gx = 0.0, gy = 0.0, gz = 0.0;
#pragma unroll
                for (int l = 0; l < H; l++) {
                    gx += g.a[0][l] * (q[H + 1 + l] - q[H - 1 - l]);
                    gy += g.a[1][l] * (sa[tj + H + 1 + l][tk + H] - sa[tj + H - 1 - l][tk + H]);
                    gz += g.a[2][l] * (sa[tj + H][tk + H + 1 + l] - sa[tj + H][tk + H - 1 - l]);
                }
                gz *= zi;
                __builtin_nontemporal_store(f0 - sc * gx, dst + q0);
                __builtin_nontemporal_store(f1 - sc * gy, dst + g.np + q0);
                __builtin_nontemporal_store(f2 - sc * gz, dst + 2 * g.np + q0);
            }
        }
        __syncthreads();
        if (i + 1 < i1) {
#pragma unroll
            for (int o = 0; o < 2 * H; o++) q[o] = q[o + 1];
            q[2 * H] = (DIV && WALLS && Nxt.fq && ox ? -1.0 : 1.0) * Nxt.q;
        }
    };
    Pf A, B;
    A.f[0] = A.f[1] = A.f[2] = 0.0;
    B.f[0] = B.f[1] = B.f[2] = 0.0;
    load_pf(A, i0, false);
    if (i0 + 1 < i1) load_pf(B, i0 + 1, true);
    for (int i = i0; i < i1; i += 2) {
        body(i, A, B);
        if (i + 1 < i1) body(i + 1, B, A);
    }
}

template <int H, bool DIV>
int launch_proj_stream_h(nova_ctx *ctx, const double *src, double sc, double *dst, int B, const BatchArgs &ba)
{
    const Grid &g = ctx->g;
    RowTiles t;
    t.nkt = (g.s[2] + kPK - 1) / kPK;
    t.njt = (g.s[1] + kPJ - 1) / kPJ;
    int64_t per = (int64_t)g.s[0] * t.nkt * t.njt / 4096;
    int ichunk = (int)(per < 8 ? 8 : (per > 64 ? 64 : per));
    t.ichunk = ichunk > g.s[0] ? g.s[0] : ichunk;
    t.nic = (g.s[0] + t.ichunk - 1) / t.ichunk;
    const dim3 grid((unsigned)(t.nkt * t.njt * t.nic), (unsigned)B);
    const bool walls = g.wall[0] || g.wall[1] || g.wall[2];
    if (walls)
        hipLaunchKernelGGL((k_proj_stream<H, DIV, true>), grid, dim3(kPK * kPJ), 0, ctx->stream, g, t, src, sc, dst,
                           ba);
    else
        hipLaunchKernelGGL((k_proj_stream<H, DIV, false>), grid, dim3(kPK * kPJ), 0, ctx->stream, g, t, src, sc, dst,
                           ba);
    return check_launch(DIV ? "div_stream" : "correct_stream");
}

int launch_div_rows(nova_ctx *ctx, const double *F, double scale, double *out, int B, const BatchArgs &ba);
int launch_correct_rows(nova_ctx *ctx, const double *p, double s, double *F, int B, const BatchArgs &ba);

template <bool DIV>
int launch_proj_stream(nova_ctx *ctx, const double *src, double sc, double *dst, int B, const BatchArgs &ba)
{
    switch (ctx->g.h) {
    case 1: return launch_proj_stream_h<1, DIV>(ctx, src, sc, dst, B, ba);
    case 2: return launch_proj_stream_h<2, DIV>(ctx, src, sc, dst, B, ba);
    case 3: return launch_proj_stream_h<3, DIV>(ctx, src, sc, dst, B, ba);
    }
    // m = 8: the halo ring exceeds one slot per thread; the row kernels serve
    return DIV ? launch_div_rows(ctx, src, sc, dst, B, ba) : launch_correct_rows(ctx, src, sc, dst, B, ba);
}

RowTiles row_tiles(const Grid &g, int ichunk)
{
    RowTiles t;
    t.nkt = (g.s[2] + kRK - 1) / kRK;
    t.njt = (g.s[1] + kRJ - 1) / kRJ;
    if (ichunk <= 0) {
        // light kernels: about 8192 blocks, chunks of 8..32 planes
        int64_t per = (int64_t)g.s[0] * t.nkt * t.njt / 8192;
        ichunk = (int)(per < 8 ? 8 : (per > 32 ? 32 : per));
    }
    t.ichunk = ichunk > g.s[0] ? g.s[0] : ichunk;
    t.nic = (g.s[0] + t.ichunk - 1) / t.ichunk;
    return t;
}

template <int H>
void launch_div_rows_h(hipStream_t st, const Grid &g, const RowTiles &t, const double *F, double scale, double *out,
                       int B, const BatchArgs &ba)
{
    const dim3 grid((unsigned)(t.nkt * t.njt * t.nic), (unsigned)B);
    const bool walls = g.wall[0] || g.wall[1] || g.wall[2];
    if (walls) hipLaunchKernelGGL((k_div_rows<H, true>), grid, dim3(kRK, kRJ), 0, st, g, t, F, scale, out, ba);
    else hipLaunchKernelGGL((k_div_rows<H, false>), grid, dim3(kRK, kRJ), 0, st, g, t, F, scale, out, ba);
}

template <int H>
void launch_correct_rows_h(hipStream_t st, const Grid &g, const RowTiles &t, const double *p, double s, double *F,
                           int B, const BatchArgs &ba)
{
    const dim3 grid((unsigned)(t.nkt * t.njt * t.nic), (unsigned)B);
    const bool walls = g.wall[0] || g.wall[1] || g.wall[2];
    if (walls) hipLaunchKernelGGL((k_correct_rows<H, true>), grid, dim3(kRK, kRJ), 0, st, g, t, p, s, F, ba);
    else hipLaunchKernelGGL((k_correct_rows<H, false>), grid, dim3(kRK, kRJ), 0, st, g, t, p, s, F, ba);
}

int launch_div_rows(nova_ctx *ctx, const double *F, double scale, double *out, int B = 1,
                    const BatchArgs &ba = BatchArgs())
{
    const RowTiles t = row_tiles(ctx->g, 0);
    switch (ctx->g.h) {
    case 1: launch_div_rows_h<1>(ctx->stream, ctx->g, t, F, scale, out, B, ba); break;
    case 2: launch_div_rows_h<2>(ctx->stream, ctx->g, t, F, scale, out, B, ba); break;
    case 3: launch_div_rows_h<3>(ctx->stream, ctx->g, t, F, scale, out, B, ba); break;
    default: launch_div_rows_h<4>(ctx->stream, ctx->g, t, F, scale, out, B, ba); break;
    }
    return check_launch("div_rows");
}

int launch_correct_rows(nova_ctx *ctx, const double *p, double s, double *F, int B = 1,
                        const BatchArgs &ba = BatchArgs())
{
    const RowTiles t = row_tiles(ctx->g, 0);
    switch (ctx->g.h) {
    case 1: launch_correct_rows_h<1>(ctx->stream, ctx->g, t, p, s, F, B, ba); break;
    case 2: launch_correct_rows_h<2>(ctx->stream, ctx->g, t, p, s, F, B, ba); break;
    case 3: launch_correct_rows_h<3>(ctx->stream, ctx->g, t, p, s, F, B, ba); break;
    default: launch_correct_rows_h<4>(ctx->stream, ctx->g, t, p, s, F, B, ba); break;
    }
    return check_launch("correct_rows");
}

// Launchers over the supported transform lengths.
#define NOVA_FFT_LENGTHS(X) X(8) X(12) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512) X(768)

bool fft_length_supported(int n)
{
    switch (n) {
#define CASE(L) case L:
        NOVA_FFT_LENGTHS(CASE)
#undef CASE
        return true;
    }
    return false;
}

int launch_fwd_z(hipStream_t st, int n, dim3 grid, const Grid &g, const double2 *tw, const int *perm,
                 const double *src, int vec, double scale, double2 *H, int Nx, int sy, bool walls, int64_t src_bs)
{
    switch (n) {
#define CASE(L)                                                                                                    \
    case L:                                                                                                        \
        if (walls) hipLaunchKernelGGL((k_fwd_z<L, true>), grid, dim3(zthreads_fwd<L>()), 0, st, g, tw, perm, src,  \
                                      vec, scale, H, Nx, sy, src_bs);                                              \
        else hipLaunchKernelGGL((k_fwd_z<L, false>), grid, dim3(zthreads_fwd<L>()), 0, st, g, tw, perm, src, vec,  \
                                scale, H, Nx, sy, src_bs);                                                         \
        return NOVA_OK;
        NOVA_FFT_LENGTHS(CASE)
#undef CASE
    }
    return fail(NOVA_ERR_UNSUPPORTED, "FFT length");
}

int launch_fft_y(hipStream_t st, int n, bool ext, bool inv, const double2 *tw, const int *perm, double2 *H,
                 int64_t nlines, int sy)
{
    switch (n) {
#define CASE(L)                                                                                   \
    case L: {                                                                                     \
        const unsigned blocks = (unsigned)((nlines + ylines<L>() - 1) / ylines<L>());             \
        if (ext && inv) hipLaunchKernelGGL((k_fft_y<L, true, true>), dim3(blocks), dim3(kLineThreads), 0, st, tw, perm, H, nlines, sy);   \
        else if (ext) hipLaunchKernelGGL((k_fft_y<L, true, false>), dim3(blocks), dim3(kLineThreads), 0, st, tw, perm, H, nlines, sy);   \
        else if (inv) hipLaunchKernelGGL((k_fft_y<L, false, true>), dim3(blocks), dim3(kLineThreads), 0, st, tw, perm, H, nlines, sy);   \
        else hipLaunchKernelGGL((k_fft_y<L, false, false>), dim3(blocks), dim3(kLineThreads), 0, st, tw, perm, H, nlines, sy);          \
        return NOVA_OK;                                                                           \
    }
        NOVA_FFT_LENGTHS(CASE)
#undef CASE
    }
    return fail(NOVA_ERR_UNSUPPORTED, "FFT length");
}

int launch_fft_x(hipStream_t st, int n, int Kc, int B, const double2 *tw, const int *perm, double2 *H, int sy,
                 const double *kpx, const double *kpy, const double *kpz)
{
    switch (n) {
#define CASE(L)                                                                                           \
    case L: {                                                                                             \
        dim3 grid((sy + xbatch_t<L>() - 1) / xbatch_t<L>(), Kc * B);                                       \
        hipLaunchKernelGGL((k_fft_x<L>), grid, dim3(kLineThreads), 0, st, tw, perm, H, sy, kpx, kpy, kpz, Kc); \
        return NOVA_OK;                                                                                   \
    }
        NOVA_FFT_LENGTHS(CASE)
#undef CASE
    }
    return fail(NOVA_ERR_UNSUPPORTED, "FFT length");
}

int launch_inv_z(hipStream_t st, int n, dim3 grid, const Grid &g, const double2 *tw, const int *perm,
                 const double2 *H, double *p, int Nx, int sy, double inv_norm, int64_t p_bs)
{
    switch (n) {
#define CASE(L)                                                                                                 \
    case L:                                                                                                     \
        hipLaunchKernelGGL((k_inv_z<L>), grid, dim3(kFftThreads), 0, st, g, tw, perm, H, p, Nx, sy, inv_norm, p_bs); \
        return NOVA_OK;
        NOVA_FFT_LENGTHS(CASE)
#undef CASE
    }
    return fail(NOVA_ERR_UNSUPPORTED, "FFT length");
}

int zlines_rt(int n)
{
    switch (n) {
#define CASE(L) case L: return zlines<L>();
        NOVA_FFT_LENGTHS(CASE)
#undef CASE
    }
    return 16;
}

int zlines_fwd_rt(int n)
{
    switch (n) {
#define CASE(L) case L: return zlines_fwd<L>();
        NOVA_FFT_LENGTHS(CASE)
#undef CASE
    }
    return 16;
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



// poisson::wavenumbersFFT_ modified wavenumber (poisson.cpp:365-389).
double kmod(double kk, double h, const double *coef, int hm)
{
    double kp = 0.0;
    if (kk > kEps)
        for (int ll = 0; ll < hm; ll++) kp += kk * coef[ll] / (ll + 1) * std::sin((ll + 1) * kk * h) / (kk * h);
    return kp;
}

}  // namespace

int poisson_create(nova_ctx *ctx, nova_poisson **out, int batch_cap)
{
    const Grid &g = ctx->g;
    if (batch_cap < 1) return fail(NOVA_ERR_ARG, "batch capacity must be >= 1");
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
    ps->batch_cap = batch_cap;
    int rc;
    if ((rc = make_len(ps->fx, ps->Nx)) || (rc = make_len(ps->fy, ps->Ty)) || (rc = make_len(ps->fz, ps->Nz))) {
        free_len(ps->fx); free_len(ps->fy); free_len(ps->fz);
        delete ps;
        return rc;
    }
    if (!fft_length_supported(ps->Nz) || !fft_length_supported(ps->Nx) || !fft_length_supported(ps->Ty)) {
        free_len(ps->fx); free_len(ps->fy); free_len(ps->fz);
        delete ps;
        return fail(NOVA_ERR_UNSUPPORTED, "FFT lengths must be one of 8..768 (2^a 3^b, listed in nova_fft.hip)");
    }
    const size_t nh = (size_t)ps->Kc * ps->Nx * ps->sy;
    NOVA_HIP(hipMalloc((void **)&ps->H, sizeof(double2) * nh * batch_cap));
    NOVA_HIP(hipMalloc((void **)&ps->dtmp, sizeof(double) * g.np * batch_cap));
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

// Select the derivative scheme: the modified wavenumbers k' of the solve
// become those of the compact first derivative, k' h = (a sin kh +
// b/2 sin 2kh) / (1 + 2 alpha cos kh), so that the solve inverts (compact
// div)(compact grad) exactly.
int poisson_set_scheme(nova_poisson *ps, int scheme)
{
    nova_ctx *ctx = ps->ctx;
    if (scheme != NOVA_SCHEME_EXPLICIT && scheme != NOVA_SCHEME_COMPACT6) return fail(NOVA_ERR_ARG, "unknown scheme");
    if (scheme == NOVA_SCHEME_COMPACT6)
        if (int rc = compact_supported(ctx)) return rc;
    const Grid &g = ctx->g;
    const double pi = 3.14159265358979323846;
    std::vector<double> kx(ps->Nx), ky(ps->sy), kz(ps->Kc);
    const double Ly = ps->wall_y ? 2.0 * ctx->desc.L[1] : ctx->desc.L[1];
    auto km = [&](double k, int d) {
        return scheme == NOVA_SCHEME_COMPACT6 ? compact_kmod(k * ctx->d[d], 1) / ctx->d[d]
                                              : kmod(k, ctx->d[d], ctx->coef, g.h);
    };
    for (int i = 0; i < ps->Nx; i++)
        kx[i] = km((2 * i < ps->Nx) ? 2.0 * pi * i / ctx->desc.L[0] : 2.0 * pi * (ps->Nx - i) / ctx->desc.L[0], 0);
    for (int j = 0; j < ps->sy; j++)
        ky[j] = km((2 * j < ps->Ty) ? 2.0 * pi * j / Ly : 2.0 * pi * (ps->Ty - j) / Ly, 1);
    for (int k = 0; k < ps->Kc; k++) kz[k] = km(2.0 * pi * k / ctx->desc.L[2], 2);
    NOVA_HIP(hipMemcpy(ps->kpx, kx.data(), sizeof(double) * ps->Nx, hipMemcpyHostToDevice));
    NOVA_HIP(hipMemcpy(ps->kpy, ky.data(), sizeof(double) * ps->sy, hipMemcpyHostToDevice));
    NOVA_HIP(hipMemcpy(ps->kpz, kz.data(), sizeof(double) * ps->Kc, hipMemcpyHostToDevice));
    ps->scheme = scheme;
    return NOVA_OK;
}

int poisson_scheme(nova_poisson *ps) { return ps->scheme; }

// Even extension (DCT-I) or the reference's periodic solve over the first
// N_y nodes of a walled y: the y transform length, its tables, the
// normalisation and the wavenumbers follow.
int poisson_set_wall_solve(nova_poisson *ps, int mode)
{
    nova_ctx *ctx = ps->ctx;
    if (mode != NOVA_WALL_SOLVE_EVEN && mode != NOVA_WALL_SOLVE_REFERENCE) return fail(NOVA_ERR_ARG, "unknown wall solve");
    if (!ctx->g.wall[1]) return fail(NOVA_ERR_ARG, "the grid has no walls in y");
    const int ext = mode == NOVA_WALL_SOLVE_EVEN;
    if (ext == ps->wall_y) return NOVA_OK;
    const int Ty = ext ? 2 * ps->Ny : ps->Ny;
    if (!fft_length_supported(Ty)) return fail(NOVA_ERR_UNSUPPORTED, "FFT length of the y solve");
    NOVA_HIP(hipDeviceSynchronize());
    nova::FftLen fy;
    if (int rc = make_len(fy, Ty)) return rc;
    free_len(ps->fy);
    ps->fy = fy;
    ps->wall_y = ext;
    ps->Ty = Ty;
    ps->norm = (double)((long)ps->Nx * (long)ps->Ty * (long)ps->Nz);
    ctx->tune_epoch++;  // recorded step graphs hold the old transform
    return poisson_set_scheme(ps, ps->scheme);  // the wavenumbers of the new y period
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
    if (ps->dtmp) (void)hipFree(ps->dtmp);
    delete ps;
    return NOVA_OK;
}

// rhs: scalar (src_is_vector = 0) or the vector whose divergence*scale is
// the right-hand side.
// The transform pipeline over B same-grid members: member b's source is
// src + b src_bs (a vector field whose divergence times its scale group's
// factor is taken, or a scalar), its potential p + b ba.p_bs.
static int run_pipeline(nova_poisson *ps, const double *src, int src_is_vector, double scale, double *p, int B,
                        const BatchArgs &ba)
{
    nova_ctx *ctx = ps->ctx;
    const Grid &g = ctx->g;
    hipStream_t st = ctx->stream;
    const bool walls = g.wall[0] || g.wall[1] || g.wall[2];
    const int zl = zlines_rt(ps->Nz), zlf = zlines_fwd_rt(ps->Nz);
    const dim3 zgrid((ps->sy + zl - 1) / zl, ps->Nx, B), zgrid_fwd((ps->sy + zlf - 1) / zlf, ps->Nx, B);
    const int64_t ylines_total = (int64_t)ps->Kc * ps->Nx * B;
    int64_t src_bs = ba.f_bs;
    int rc;
    if (src_is_vector && (ctx->proj_variant >= 1 || B > 1)) {
        // divergence by a streamed kernel into a scalar, then a plain r2c
        {
            TimedScope ts(ctx, "div_rows");
            rc = ctx->proj_variant == 2 ? launch_proj_stream<true>(ctx, src, scale, ps->dtmp, B, ba)
                                        : launch_div_rows(ctx, src, scale, ps->dtmp, B, ba);
        }
        if (rc) return rc;
        src = ps->dtmp;
        src_is_vector = 0;
        scale = 1.0;
        src_bs = g.np;
    }
    {
        TimedScope ts(ctx, src_is_vector ? "fft_div_r2c_z" : "fft_r2c_z");
        rc = launch_fwd_z(st, ps->Nz, zgrid_fwd, g, ps->fz.tw, ps->fz.perm, src, src_is_vector, scale, ps->H, ps->Nx,
                          ps->sy, walls, src_bs);
    }
    if (rc || (rc = check_launch("poisson fwd_z"))) return rc;
    {
        TimedScope ts(ctx, "fft_y");
        rc = launch_fft_y(st, ps->Ty, ps->wall_y, false, ps->fy.tw, ps->fy.perm, ps->H, ylines_total, ps->sy);
    }
    if (rc || (rc = check_launch("poisson fft_y"))) return rc;
    {
        TimedScope ts(ctx, "fft_x_solve");
        rc = launch_fft_x(st, ps->Nx, ps->Kc, B, ps->fx.tw, ps->fx.perm, ps->H, ps->sy, ps->kpx, ps->kpy, ps->kpz);
    }
    if (rc || (rc = check_launch("poisson fft_x"))) return rc;
    {
        TimedScope ts(ctx, "fft_y");
        rc = launch_fft_y(st, ps->Ty, ps->wall_y, true, ps->fy.tw, ps->fy.perm, ps->H, ylines_total, ps->sy);
    }
    if (rc || (rc = check_launch("poisson ifft_y"))) return rc;
    {
        TimedScope ts(ctx, "fft_c2r_z");
        rc = launch_inv_z(st, ps->Nz, zgrid, g, ps->fz.tw, ps->fz.perm, ps->H, p, ps->Nx, ps->sy, 1.0 / ps->norm,
                          ba.p_bs);
    }
    if (rc || (rc = check_launch("poisson inv_z"))) return rc;
    return NOVA_OK;
}

int poisson_run(nova_poisson *ps, const double *src, int src_is_vector, double scale, double *p)
{
    return run_pipeline(ps, src, src_is_vector, scale, p, 1, BatchArgs());
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
    if (ps->scheme == NOVA_SCHEME_COMPACT6) {
        // compact divergence passes, the transform pipeline on the scalar,
        // compact gradient-correction passes
        if (int rc = compact_div(ctx, F, 1.0 / s, ps->dtmp)) return rc;
        if (int rc = run_pipeline(ps, ps->dtmp, 0, 1.0, p, 1, BatchArgs())) return rc;
        return compact_grad_correct(ctx, p, s, F);
    }
    if (int rc = poisson_run(ps, F, 1, 1.0 / s, p)) return rc;
    // the reference's fills: p.correctBoundaryConditions() before the
    // gradient (app/shearLayer/UEqn.H:32-34)
    if (int rc = wall_fill(ctx, p, 1)) return rc;
    const Grid &g = ctx->g;
    dim3 grid((g.s[2] + 63) / 64, (g.s[1] + 3) / 4, g.s[0]);
    TimedScope ts(ctx, "grad_correct");
    if (ctx->proj_variant == 2) return launch_proj_stream<false>(ctx, p, s, F, 1, BatchArgs());
    if (ctx->proj_variant == 1) return launch_correct_rows(ctx, p, s, F);
    if (g.wall[0] || g.wall[1] || g.wall[2])
        hipLaunchKernelGGL(k_correct<true>, grid, dim3(64, 4, 1), 0, ctx->stream, g, p, s, F);
    else
        hipLaunchKernelGGL(k_correct<false>, grid, dim3(64, 4, 1), 0, ctx->stream, g, p, s, F);
    return check_launch("project correct");
}

// Projection of B same-grid fields F + b f_bs (potentials p + b p_bs):
// members b < split use s_lo, the others s_hi.  Always the streamed
// divergence / correction pipeline.
int project_batch(nova_poisson *ps, double *F, int64_t f_bs, int B, int split, double s_lo, double s_hi, double *p,
                  int64_t p_bs)
{
    nova_ctx *ctx = ps->ctx;
    if (B < 1 || B > ps->batch_cap) return fail(NOVA_ERR_ARG, "batch exceeds the Poisson batch capacity");
    if (ps->scheme != NOVA_SCHEME_EXPLICIT) return fail(NOVA_ERR_UNSUPPORTED, "batched projections are explicit");
    BatchArgs ba;
    ba.f_bs = f_bs;
    ba.p_bs = p_bs;
    ba.split = split;
    ba.scale_hi = 1.0 / s_hi;
    ba.s_hi = s_hi;
    if (int rc = run_pipeline(ps, F, 1, 1.0 / s_lo, p, B, ba)) return rc;
    TimedScope ts(ctx, "grad_correct");
    if (ctx->proj_variant == 2) return launch_proj_stream<false>(ctx, p, s_lo, F, B, ba);
    return launch_correct_rows(ctx, p, s_lo, F, B, ba);
}

}  // namespace nova

namespace nova {
nova_ctx *poisson_ctx(nova_poisson *ps) { return ps->ctx; }
}  // namespace nova

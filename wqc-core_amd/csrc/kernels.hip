// kernels.hip — gfx950 kernels of the MPS engine (FP64 complex, wave64).
//
// All kernels take a device job table so that every independent job of a brickwork
// layer (up to ~50 two-site splits at 100 qubits) runs in one launch: grid.y (or grid.z)
// indexes the job, and jobs smaller than the launch's largest exit early.
#include "kernels.h"

#include <math.h>

namespace wqc {
namespace {

__device__ __forceinline__ double2 cz() { return make_double2(0.0, 0.0); }
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double2 cscale(double2 a, double k) { return make_double2(a.x * k, a.y * k); }
// acc += a * b
__device__ __forceinline__ void cmac(double2& acc, double2 a, double2 b) {
    acc.x = fma(a.x, b.x, acc.x);
    acc.x = fma(-a.y, b.y, acc.x);
    acc.y = fma(a.x, b.y, acc.y);
    acc.y = fma(a.y, b.x, acc.y);
}
// acc += a * conj(b)
__device__ __forceinline__ void cmacc(double2& acc, double2 a, double2 b) {
    acc.x = fma(a.x, b.x, acc.x);
    acc.x = fma(a.y, b.y, acc.x);
    acc.y = fma(a.y, b.x, acc.y);
    acc.y = fma(-a.x, b.y, acc.y);
}
// acc += conj(a) * b
__device__ __forceinline__ void ccmac(double2& acc, double2 a, double2 b) {
    acc.x = fma(a.x, b.x, acc.x);
    acc.x = fma(a.y, b.y, acc.x);
    acc.y = fma(a.x, b.y, acc.y);
    acc.y = fma(-a.y, b.x, acc.y);
}

typedef double d4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ d4v mfma_f64(double a, double b, d4v c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// ------------------------------------------------------------------ one-qubit update
__global__ __launch_bounds__(256) void k_one_qubit(const OneQubitJob* __restrict__ jobs) {
    const OneQubitJob jb = jobs[blockIdx.y];
    const long total = (long)jb.l * jb.r;
    const double2 u00 = jb.u[0], u01 = jb.u[1], u10 = jb.u[2], u11 = jb.u[3];
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (long)gridDim.x * blockDim.x) {
        const long a = e / jb.r, b = e - a * jb.r;
        double2* p0 = jb.site + a * 2 * jb.r + b;
        double2* p1 = p0 + jb.r;
        const double2 v0 = *p0, v1 = *p1;
        double2 o0 = cz(), o1 = cz();
        cmac(o0, u00, v0);
        cmac(o0, u01, v1);
        cmac(o1, u10, v0);
        cmac(o1, u11, v1);
        *p0 = o0;
        *p1 = o1;
    }
}

// ------------------------------------------------------------------ merge + U
// theta' = U4 (L R) on FP64 MFMA (v_mfma_f64_16x16x4f64, four real products per complex one:
// Re = Lr Rr - Li Ri, Im = Lr Ri + Li Rr, so every entry carries the rounding of a plain complex
// product; exact zeros of the spectrum stay at that noise floor). A workgroup owns 32
// values of a and 32 of b: theta rows {2a, 2a+1} and columns {b, dr + b}, the four entries (s, t)
// that the two-qubit unitary mixes. K advances in steps of 16 staged through LDS as real /
// imaginary planes (L rows ordered a*2 + s, R columns t*32 + b). Wave w holds the 16 x 16 (a, b)
// quarter (16 (w >> 1), 16 (w & 1)) for all four (s, t), so the unitary is applied in registers
// and theta' (or its conjugate transpose) is written directly.
constexpr int MA = 32, MK = 16;

__global__ __launch_bounds__(256, 2) void k_merge(const MergeJob* __restrict__ jobs) {
    const MergeJob jb = jobs[blockIdx.z];
    const int a0 = blockIdx.y * MA, b0 = blockIdx.x * MA;
    if (a0 >= jb.dl || b0 >= jb.dr) return;
    __shared__ double Lr[2 * MA][MK + 1], Li[2 * MA][MK + 1], Rr[MK][2 * MA + 1], Ri[MK][2 * MA + 1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int aw = (wave >> 1) * 16, bw = (wave & 1) * 16;
    const int li = lane & 15, lk = lane >> 4;
    const long ldl = jb.bond, ldr = 2L * jb.dr;
    d4v e1[2][2], e2[2][2], e3[2][2], e4[2][2];  // [s][t]: Lr Rr, Li Ri, Lr Ri, Li Rr
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            e1[s][t] = d4v{0, 0, 0, 0};
            e2[s][t] = e1[s][t];
            e3[s][t] = e1[s][t];
            e4[s][t] = e1[s][t];
        }
    // Each thread stages 4 entries of L and 4 of R per K tile; the next tile is fetched into
    // registers while the MFMAs consume the current one.
    constexpr int PER = (2 * MA * MK) / 256;
    double2 rl[PER], rr[PER];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int e = tid + 256 * u;
            {
                const int lr = e / MK, kk = e % MK;
                const int a = a0 + (lr >> 1), s = lr & 1, k = k0 + kk;
                rl[u] = (a < jb.dl && k < jb.bond) ? jb.L[(long)(2 * a + s) * ldl + k] : cz();
            }
            {
                const int kk = e / (2 * MA), lc = e % (2 * MA);
                const int t = lc / MA, b = b0 + (lc % MA), k = k0 + kk;
                rr[u] = (b < jb.dr && k < jb.bond) ? jb.R[(long)k * ldr + (long)t * jb.dr + b] : cz();
            }
        }
    };
    fetch(0);
    for (int k0 = 0; k0 < jb.bond; k0 += MK) {
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int e = tid + 256 * u;
            Lr[e / MK][e % MK] = rl[u].x;
            Li[e / MK][e % MK] = rl[u].y;
            Rr[e / (2 * MA)][e % (2 * MA)] = rr[u].x;
            Ri[e / (2 * MA)][e % (2 * MA)] = rr[u].y;
        }
        __syncthreads();
        if (k0 + MK < jb.bond) fetch(k0 + MK);
#pragma unroll
        for (int ks = 0; ks < MK / 4; ++ks) {
            const int kk = 4 * ks + lk;
            double ar[2], ai[2], br[2], bi[2];
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                ar[s] = Lr[(aw + li) * 2 + s][kk];
                ai[s] = Li[(aw + li) * 2 + s][kk];
            }
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                br[t] = Rr[kk][t * MA + bw + li];
                bi[t] = Ri[kk][t * MA + bw + li];
            }
#pragma unroll
            for (int s = 0; s < 2; ++s)
#pragma unroll
                for (int t = 0; t < 2; ++t) {
                    e1[s][t] = mfma_f64(ar[s], br[t], e1[s][t]);
                    e2[s][t] = mfma_f64(ai[s], bi[t], e2[s][t]);
                    e3[s][t] = mfma_f64(ar[s], bi[t], e3[s][t]);
                    e4[s][t] = mfma_f64(ai[s], br[t], e4[s][t]);
                }
        }
        __syncthreads();
    }
    // MFMA f64 output: element j of lane l is row 4j + l/16 (a), column l%16 (b) of the quarter.
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int a = a0 + aw + 4 * j + lk, b = b0 + bw + li;
        if (a >= jb.dl || b >= jb.dr) continue;
        double2 th[4];  // index s*2 + t
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int t = 0; t < 2; ++t)
                th[s * 2 + t] = make_double2(e1[s][t][j] - e2[s][t][j], e3[s][t][j] + e4[s][t][j]);
#pragma unroll
        for (int so = 0; so < 2; ++so)
#pragma unroll
            for (int to = 0; to < 2; ++to) {
                double2 v = cz();
#pragma unroll
                for (int i = 0; i < 4; ++i) cmac(v, jb.u4[(so * 2 + to) * 4 + i], th[i]);
                const long row = 2L * a + so, col = (long)to * jb.dr + b;
                if (!jb.transpose)
                    jb.out[row * ldr + col] = v;
                else
                    jb.out[col * (2L * jb.dl) + row] = cconj(v);
            }
    }
}

// ------------------------------------------------------------------ sums of squares
__global__ __launch_bounds__(256) void k_sumsq(const double2* const* __restrict__ ptrs,
                                               const long* __restrict__ counts,
                                               double* __restrict__ partial, int nblk) {
    const int job = blockIdx.y;
    const double2* p = ptrs[job];
    const long n = counts[job];
    double acc = 0.0;
    for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < n; e += (long)nblk * 256) {
        const double2 v = p[e];
        acc = fma(v.x, v.x, acc);
        acc = fma(v.y, v.y, acc);
    }
    acc = wave_sum(acc);
    __shared__ double ws[4];
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) partial[(long)job * nblk + blockIdx.x] = ((ws[0] + ws[1]) + ws[2]) + ws[3];
}

__global__ void k_sumsq_final(const double* __restrict__ partial, double* __restrict__ out,
                              int njobs, int nblk) {
    const int job = blockIdx.x * blockDim.x + threadIdx.x;
    if (job >= njobs) return;
    double s = 0.0;
    for (int i = 0; i < nblk; ++i) s += partial[(long)job * nblk + i];
    out[job] = s;
}

// checked_svd's dust rule: entries with |z| < rel * ||A||_F become exactly zero.
__global__ __launch_bounds__(256) void k_dust(double2* const* __restrict__ ptrs,
                                              const long* __restrict__ counts,
                                              const double* __restrict__ fro2, double rel) {
    const int job = blockIdx.y;
    double2* p = ptrs[job];
    const long n = counts[job];
    const double dust = sqrt(fro2[job]) * rel;
    for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < n; e += (long)gridDim.x * 256) {
        const double2 v = p[e];
        if (hypot(v.x, v.y) < dust) p[e] = cz();
    }
}

__global__ __launch_bounds__(256) void k_identity(double2* const* __restrict__ ptrs,
                                                  const int* __restrict__ dims) {
    const int job = blockIdx.y;
    double2* p = ptrs[job];
    const long n = dims[job];
    for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < n * n; e += (long)gridDim.x * 256) {
        const long i = e / n, j = e - i * n;
        p[e] = make_double2(i == j ? 1.0 : 0.0, 0.0);
    }
}

// ------------------------------------------------------------------ block Jacobi round
// One workgroup (4 waves) orthogonalizes the 2*JB rows of one block pair of one job:
//  1. Gram matrix G = P P^H of the pair on FP64 MFMA (v_mfma_f64_16x16x4f64), operands
//     loaded from HBM straight into registers, the columns split over the 4 waves;
//  2. Hermitian eigensolve of G by parallel cyclic Jacobi in LDS: per step 16 disjoint
//     rotations, each thread owning one 2x2 block of G and of the accumulated rotation T
//     (G <- M G M^H and T <- M T in one phase);
//  3. P <- T P over the columns of X and of J, again on FP64 MFMA with T in registers.
// A pair (i, j) needs a rotation when |G_ij| > tol sqrt(G_ii G_jj) (relative, as the
// reference's Jacobi, src/tn/svd_guard.rs:178), tol = tol_base sqrt(m) per job: small
// singular values come out accurate relative to themselves, so the rank count above 1e-12
// and the dust rule see the same values an exact SVD gives. Rows of squared norm <= nu2 =
// nu2_base / n (the host passes 1e-26, so that all such rows of a job together hold at most
// 1e-13 of singular value, an order below the 1e-12 cutoff) take part in no rotation: they cannot move a kept value or a decision,
// and orthogonalizing rounding noise against itself would never settle. The in-block eigensolve runs at most JINNER sweeps; the
// next outer round recomputes the pair's Gram matrix from the rows themselves, so an
// unfinished inner solve only costs another round. rotated[job] records that the sweep
// is not converged; stats[0..1] count inner sweeps and pair solves.
constexpr int JB = 16;
constexpr int JP = 2 * JB;

// De Rijk-style ordering of a pair step: after the in-block eigensolve, the rotated rows are
// ranked by their Gram diagonal (descending; ties by row index) and the k-th largest is
// written to the k-th lowest row index of the pair. Every pair step then agrees on one
// global order (large rows first), which removes the slow first phase of block Jacobi with
// unordered eigenvectors (about half the sweeps on decaying spectra). Slots beyond the
// matrix (s_row < 0) keep their place. Threads 0..JP-1 write dst[k], the destination slot
// of local row k.
template <typename Key>
__device__ __forceinline__ void pair_order(int tid, const int* s_row, const Key* diag, int stride,
                                           int* dst) {
    if (tid >= JP) return;
    const int rk = s_row[tid];
    int out = tid;
    if (rk >= 0) {
        const Key key = diag[tid * stride];
        int rank = 0, na = 0, nb = 0;
        for (int j = 0; j < JP; ++j) {
            const int rj = s_row[j];
            if (rj < 0) continue;
            if (j < JB)
                ++na;
            else
                ++nb;
            if (j == tid) continue;
            const Key kj = diag[j * stride];
            if (kj > key || (kj == key && rj < rk)) ++rank;
        }
        // The slot holding the rank-th smallest row index. Each half's valid rows are a prefix
        // of it in ascending order, and the half of the lower block comes first.
        const bool a_first = (na == 0 || nb == 0) ? na > 0 : s_row[0] < s_row[JB];
        if (a_first)
            out = rank < na ? rank : JB + rank - na;
        else
            out = rank < nb ? JB + rank : rank - nb;
    }
    dst[tid] = out;
}


// Reciprocal and reciprocal square root from the hardware estimates plus two Newton steps
// (full FP64 precision), a third of the dependent instructions of IEEE division and sqrt.
__device__ __forceinline__ double rcp_nr(double x) {
    double y = __builtin_amdgcn_rcp(x);
    double e = fma(-x, y, 1.0);
    y = fma(y, e, y);
    e = fma(-x, y, 1.0);
    return fma(y, e, y);
}

__device__ __forceinline__ double rsq_nr(double x) {
    double y = __builtin_amdgcn_rsq(x);
    const double h = 0.5 * x;
    double e = fma(-h * y, y, 0.5);
    y = fma(y, e, y);
    e = fma(-h * y, y, 0.5);
    return fma(y, e, y);
}

// Rotation pair k of step `step` of the circle method on JP indices.
__device__ __forceinline__ void circle_pair(int step, int k, int& p, int& q) {
    if (k == 0) {
        p = JP - 1;
        q = step;
    } else {
        p = step + k;
        if (p >= JP - 1) p -= JP - 1;
        q = step - k;
        if (q < 0) q += JP - 1;
    }
}

// Rotation pair k of step `step` (0 .. JB-1) of a cross-only sweep: row k of the first
// block with row JB + (k + step) mod JB of the second, every cross pair once per JB steps.
__device__ __forceinline__ void cross_pair(int step, int k, int& p, int& q) {
    p = k;
    q = JB + ((k + step) & (JB - 1));
}

// Complex 2x2 Jacobi rotation M = [[cs, m01], [m10, cs]] that zeroes G_pq of the Hermitian
// block [[a, c], [conj(c), d]]: M G M^H = diag(a - tg, d + tg). act = 0 when the pair is
// orthogonal to tol (relative) or one of its rows is below nu2.
struct Rot64 {
    int act;
    double cs, tg;
    double2 m01, m10;
};

__device__ __forceinline__ Rot64 rot_params64(double a, double d, double2 c, double tol, double nu2) {
    Rot64 r;
    r.act = 0;
    r.cs = 1.0;
    r.tg = 0.0;
    r.m01 = cz();
    r.m10 = cz();
    const double g2 = fma(c.x, c.x, c.y * c.y);
    if (!(g2 > 0.0) || !(fmin(a, d) > nu2) || !(g2 > tol * tol * (a * d))) return r;
    r.act = 1;
    const double rg = rsq_nr(g2);
    const double zeta = 0.5 * (d - a) * rg;
    const double w = fma(zeta, zeta, 1.0);
    const double t = copysign(rcp_nr(fabs(zeta) + w * rsq_nr(w)), zeta);
    const double cs = rsq_nr(fma(t, t, 1.0));
    const double sn = cs * t;
    const double phx = c.x * rg, phy = c.y * rg;
    r.cs = cs;
    r.m01 = make_double2(-phx * sn, -phy * sn);
    r.m10 = make_double2(phx * sn, -phy * sn);
    r.tg = t * (g2 * rg);
    return r;
}

// Rotation k of a step is computed by lanes k, k + 16, ... of every wave (as the rotation
// of their column blocks); the row-block rotation of a lane is fetched from lane bk of its
// wave instead of being derived a second time. m10 = (-m01.x, m01.y) by construction.
__device__ __forceinline__ Rot64 rot_from_lane(const Rot64& own, int src) {
    Rot64 r;
    r.act = __shfl(own.act, src, 64);
    r.cs = __shfl(own.cs, src, 64);
    r.tg = __shfl(own.tg, src, 64);
    r.m01.x = __shfl(own.m01.x, src, 64);
    r.m01.y = __shfl(own.m01.y, src, 64);
    r.m10 = make_double2(-r.m01.x, r.m01.y);
    return r;
}


// NT = 256: four waves, three workgroups per CU (the bulk of a batch's launches). NT = 512:
// eight waves share the Gram and the rotation of one block pair, for launches with no more
// block pairs than CUs, where a round lasts as long as one workgroup; the eigensolve stays on
// the first four waves.
template <int NT>
__global__ __launch_bounds__(NT, NT == 256 ? 3 : 1) void k_jacobi_round(const JacobiJob* __restrict__ jobs,
                                                      int round, double tol_base, double nu2_base,
                                                      int inner_cap, int* __restrict__ rotated,
                                                      const int* __restrict__ done,
                                                      unsigned long long* __restrict__ stats,
                                                      unsigned long long* __restrict__ offmax) {
    constexpr int NW = NT / 64;
    const int job = blockIdx.y;
    if (done[job]) return;
    const JacobiJob jb = jobs[job];
    // Per-job tolerance tol_base sqrt(m) and null-row floor nu2_base / n (see the header comment).
    const double tol = tol_base * sqrt((double)jb.m);
    const double nu2 = nu2_base / (double)jb.n;
    const int nb = (jb.n + JB - 1) / JB;
    const int nbe = nb + (nb & 1);
    if (round >= nbe - 1) return;
    const int kpair = blockIdx.x;
    if (kpair >= nbe / 2) return;
    int bi, bj;
    if (kpair == 0) {
        bi = nbe - 1;
        bj = round;
    } else {
        bi = (round + kpair) % (nbe - 1);
        bj = (round - kpair + nbe - 1) % (nbe - 1);
    }

    __shared__ double2 G[JP][JP + 1];
    // Partial Gram tiles of two waves; afterwards T (the accumulated rotation, updated in
    // place: a step's tiles are disjoint) and the second G buffer of the eigensolve. About
    // 51 KB of LDS in all, so three workgroups share a CU.
    __shared__ __align__(16) double s_aux[NW == 8 ? 4 * 6 * 4 * 64 : 4 * JP * (JP + 1)];
    __shared__ int s_row[JP];
    __shared__ int s_any;
    __shared__ int s_intra;
    __shared__ int s_dst[JP];
    __shared__ unsigned long long s_off;
    double (*Gp)[6][4][64] = reinterpret_cast<double (*)[6][4][64]>(s_aux);
    double2 (*T)[JP + 1] = reinterpret_cast<double2 (*)[JP + 1]>(s_aux);
    double2 (*G2)[JP + 1] = reinterpret_cast<double2 (*)[JP + 1]>(s_aux + 2 * JP * (JP + 1));

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid < JP) {
        const int b = tid < JB ? bi : bj;
        const int r = b * JB + (tid & (JB - 1));
        s_row[tid] = (b < nb && r < jb.n) ? r : -1;
    }
    if (tid == 0) {
        s_any = 0;
        s_intra = 0;
        s_off = 0;
    }
    __syncthreads();

    // ---- 1. Gram on MFMA. Lane (i = lane%16, g = lane/16) holds columns c0+4g..c0+4g+3 of
    // local rows i and 16+i; MFMA step s uses column c0 + 4g + s for both operands.
    {
        const int i = lane & 15, g = lane >> 4;
        const int r0 = s_row[i], r1 = s_row[16 + i];
        const double2* x0 = r0 >= 0 ? jb.X + (long)r0 * jb.ldx : nullptr;
        const double2* x1 = r1 >= 0 ? jb.X + (long)r1 * jb.ldx : nullptr;
        // 3M products: T1 = sum x.re y.re, T2 = sum x.im y.im, T3 = sum (x.re + x.im)(y.re - y.im);
        // Re G = T1 + T2, Im G = T3 - T1 + T2. Column chunks are double-buffered in registers.
        d4v p00 = {0, 0, 0, 0}, q00 = p00, t00 = p00, p01 = p00, q01 = p00, t01 = p00, p11 = p00, q11 = p00, t11 = p00;
        double2 ua[4], va[4], ub[4], vb[4];
#define WQC_GRAM_LOAD(C0, U, V)                                  \
    _Pragma("unroll") for (int s = 0; s < 4; ++s) {              \
        const int c = (C0) + 4 * g + s;                          \
        U[s] = (x0 && c < jb.m) ? x0[c] : cz();                  \
        V[s] = (x1 && c < jb.m) ? x1[c] : cz();                  \
    }
#define WQC_GRAM_ACC(U, V)                                       \
    _Pragma("unroll") for (int s = 0; s < 4; ++s) {              \
        const double up = U[s].x + U[s].y, um = U[s].x - U[s].y; \
        const double vp = V[s].x + V[s].y, vm = V[s].x - V[s].y; \
        p00 = mfma_f64(U[s].x, U[s].x, p00);                     \
        q00 = mfma_f64(U[s].y, U[s].y, q00);                     \
        t00 = mfma_f64(up, um, t00);                             \
        p01 = mfma_f64(U[s].x, V[s].x, p01);                     \
        q01 = mfma_f64(U[s].y, V[s].y, q01);                     \
        t01 = mfma_f64(up, vm, t01);                             \
        p11 = mfma_f64(V[s].x, V[s].x, p11);                     \
        q11 = mfma_f64(V[s].y, V[s].y, q11);                     \
        t11 = mfma_f64(vp, vm, t11);                             \
    }
        int c0 = wave * 16;
        if (c0 < jb.m) WQC_GRAM_LOAD(c0, ua, va)
        while (c0 < jb.m) {
            const int c1 = c0 + 16 * NW;
            if (c1 < jb.m) WQC_GRAM_LOAD(c1, ub, vb)
            WQC_GRAM_ACC(ua, va)
            if (c1 >= jb.m) break;
            const int c2 = c1 + 16 * NW;
            if (c2 < jb.m) WQC_GRAM_LOAD(c2, ua, va)
            WQC_GRAM_ACC(ub, vb)
            c0 = c2;
        }
#undef WQC_GRAM_LOAD
#undef WQC_GRAM_ACC
        d4v a00r = p00 + q00, a00i = t00 - p00 + q00;
        d4v a01r = p01 + q01, a01i = t01 - p01 + q01;
        d4v a11r = p11 + q11, a11i = t11 - p11 + q11;
        // Deterministic reduction over the waves: (w0 + w2) + (w1 + w3).
#define WQC_GP_ALL(OP) OP(0, a00r) OP(1, a00i) OP(2, a01r) OP(3, a01i) OP(4, a11r) OP(5, a11i)
#define WQC_GP_STORE(t, v) for (int j = 0; j < 4; ++j) Gp[slot][t][j][lane] = v[j];
#define WQC_GP_ADD(t, v) for (int j = 0; j < 4; ++j) v[j] += Gp[slot][t][j][lane];
        if constexpr (NW == 8) {
            // (w0 + w4), (w1 + w5), (w2 + w6), (w3 + w7) first.
            if (wave >= 4) {
                const int slot = wave - 4;
                WQC_GP_ALL(WQC_GP_STORE)
            }
            __syncthreads();
            if (wave < 4) {
                const int slot = wave;
                WQC_GP_ALL(WQC_GP_ADD)
            }
            __syncthreads();
        }
        if (wave >= 2 && wave < 4) {
            const int slot = wave - 2;
            WQC_GP_ALL(WQC_GP_STORE)
        }
        __syncthreads();
        if (wave < 2) {
            const int slot = wave;
            WQC_GP_ALL(WQC_GP_ADD)
        }
        __syncthreads();
        if (wave == 1) {
            const int slot = 0;
            WQC_GP_ALL(WQC_GP_STORE)
        }
        __syncthreads();
#undef WQC_GP_ALL
#undef WQC_GP_STORE
#undef WQC_GP_ADD
        if (wave == 0) {
            // MFMA f64 output: element j of lane l is row 4j + l/16, column l%16.
            const int r = lane >> 4, c = lane & 15;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const double g00r = a00r[j] + Gp[0][0][j][lane], g00i = a00i[j] + Gp[0][1][j][lane];
                const double g01r = a01r[j] + Gp[0][2][j][lane], g01i = a01i[j] + Gp[0][3][j][lane];
                const double g11r = a11r[j] + Gp[0][4][j][lane], g11i = a11i[j] + Gp[0][5][j][lane];
                const int rr = 4 * j + r;
                G[rr][c] = make_double2(g00r, g00i);
                G[rr][16 + c] = make_double2(g01r, g01i);
                G[16 + c][rr] = make_double2(g01r, -g01i);
                G[16 + rr][16 + c] = make_double2(g11r, g11i);
            }
        }
    }
    __syncthreads();
    {
        bool off = false, intra = false;
        double worst = 0.0;
        for (int e = tid; e < JP * JP; e += NT) {
            const int i = e / JP, j = e % JP;
            if (i < j) {
                const double2 c = G[i][j];
                const double g2 = fma(c.x, c.x, c.y * c.y);
                const double di = G[i][i].x, dj = G[j][j].x;
                if (g2 > 0.0 && fmin(di, dj) > nu2) {
                    if (g2 > tol * tol * (di * dj)) {
                        off = true;
                        if ((i < JB) == (j < JB)) intra = true;
                    }
                    if (offmax) worst = fmax(worst, g2 / (di * dj));
                }
            }
        }
        if (off) s_any = 1;
        if (intra) s_intra = 1;
        if (offmax) {
            // The largest squared ratio over the wave first (one LDS atomic per wave), then its
            // square root: sqrt is monotonic, so this is exactly the largest relative
            // off-diagonal. Non-negative doubles order as their bit patterns.
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) worst = fmax(worst, __shfl_xor(worst, o, 64));
            if (lane == 0 && worst > 0.0) atomicMax(&s_off, (unsigned long long)__double_as_longlong(sqrt(worst)));
        }
    }
    __syncthreads();
    if (offmax && tid == 0 && s_off) atomicMax(&offmax[job], s_off);
    // Every wave reads the flag before thread 0 clears it.
    const int need = s_any;
    const int intra0 = s_intra;
    __syncthreads();
    if (!need) return;
    if (tid == 0) {
        rotated[job] = 1;
        s_any = 0;
    }
    for (int e = tid; e < JP * JP; e += NT) {
        const int i = e / JP, j = e % JP;
        T[i][j] = make_double2(i == j ? 1.0 : 0.0, 0.0);
    }
    __syncthreads();

    // ---- 2. eigensolve of G (T accumulates the row rotations). Every thread derives the
    // parameters of its own two rotations from the current buffers, so a step is one
    // read-update-write pass with G ping-ponged between two LDS buffers and a single barrier;
    // T is only read and written at the thread's own tile, so it is updated in place.
    double2 (*Gc)[JP + 1] = G;
    double2 (*const Tc)[JP + 1] = T;
    double2 (*Gn)[JP + 1] = G2;
    int sweeps_done = 0;
    const int bk = tid >> 4, bk2 = tid & 15;
    for (int sweep = 0; sweep < inner_cap; ++sweep) {
        ++sweeps_done;
        // The pair's two blocks come out of their previous rounds mutually orthogonal inside;
        // when the check found no block-internal pair above tolerance, the first sweep
        // rotates only the JB x JB cross pairs (JB steps instead of 2 JB - 1).
        const bool cross = sweep == 0 && !intra0;
        const int nsteps = cross ? JB : JP - 1;
        for (int step = 0; step < nsteps; ++step) {
            if (tid < 256) {  // the eigensolve's 16 x 16 tiles of 2 x 2
                int p1, q1, p2, q2;
                if (cross) {
                    cross_pair(step, bk, p1, q1);
                    cross_pair(step, bk2, p2, q2);
                } else {
                    circle_pair(step, bk, p1, q1);
                    circle_pair(step, bk2, p2, q2);
                }
                const Rot64 r2 = rot_params64(Gc[p2][p2].x, Gc[q2][q2].x, Gc[p2][q2], tol, nu2);
                const Rot64 r1 = rot_from_lane(r2, bk & 15);
                double2 h00 = Gc[p1][p2], h01 = Gc[p1][q2], h10 = Gc[q1][p2], h11 = Gc[q1][q2];
                if (r1.act) {
                    const double cs = r1.cs;
                    const double2 m01 = r1.m01, m10 = r1.m10;
                    const double2 x00 = cadd(cscale(h00, cs), cmul(m01, h10));
                    const double2 x01 = cadd(cscale(h01, cs), cmul(m01, h11));
                    const double2 x10 = cadd(cmul(m10, h00), cscale(h10, cs));
                    const double2 x11 = cadd(cmul(m10, h01), cscale(h11, cs));
                    h00 = x00;
                    h01 = x01;
                    h10 = x10;
                    h11 = x11;
                    // T (updated in place) changes only with the row rotation.
                    const double2 t00 = Tc[p1][p2], t01 = Tc[p1][q2], t10 = Tc[q1][p2], t11 = Tc[q1][q2];
                    Tc[p1][p2] = cadd(cscale(t00, cs), cmul(m01, t10));
                    Tc[p1][q2] = cadd(cscale(t01, cs), cmul(m01, t11));
                    Tc[q1][p2] = cadd(cmul(m10, t00), cscale(t10, cs));
                    Tc[q1][q2] = cadd(cmul(m10, t01), cscale(t11, cs));
                }
                if (r2.act) {
                    const double cs = r2.cs;
                    const double2 c01 = cconj(r2.m01), c10 = cconj(r2.m10);
                    const double2 n00 = cadd(cscale(h00, cs), cmul(h01, c01));
                    const double2 n01 = cadd(cmul(h00, c10), cscale(h01, cs));
                    const double2 n10 = cadd(cscale(h10, cs), cmul(h11, c01));
                    const double2 n11 = cadd(cmul(h10, c10), cscale(h11, cs));
                    h00 = n00;
                    h01 = n01;
                    h10 = n10;
                    h11 = n11;
                }
                if (bk == bk2 && r1.act) {
                    h00 = make_double2(Gc[p1][p1].x - r1.tg, 0.0);
                    h11 = make_double2(Gc[q1][q1].x + r1.tg, 0.0);
                    h01 = cz();
                    h10 = cz();
                    s_any = 1;
                }
                Gn[p1][p2] = h00;
                Gn[p1][q2] = h01;
                Gn[q1][p2] = h10;
                Gn[q1][q2] = h11;
            }
            __syncthreads();
            double2 (*sw)[JP + 1] = Gc;
            Gc = Gn;
            Gn = sw;
        }
        // Every wave reads the flag before thread 0 clears it for the next sweep.
        const int again = s_any;
        __syncthreads();
        if (!again || sweep + 1 >= inner_cap) break;
        // The next sweep is a no-op unless some pair passes rot_params64's test on the current
        // G in either orientation (an inactive rotation leaves G and T unchanged, so every
        // later step of the sweep sees the same entries): skip it then, with the same result.
        if (tid == 0) s_any = 0;
        __syncthreads();
        {
            bool act = false;
            for (int e = tid; e < JP * JP; e += NT) {
                const int i = e / JP, j = e % JP;
                if (i == j) continue;
                const double2 c = Gc[i][j];
                const double g2 = fma(c.x, c.x, c.y * c.y);
                const double a = Gc[i][i].x, d = Gc[j][j].x;
                if (g2 > 0.0 && fmin(a, d) > nu2 && g2 > tol * tol * (a * d)) act = true;
            }
            if (act) s_any = 1;
        }
        __syncthreads();
        const int next = s_any;
        __syncthreads();
        if (!next) break;
        if (tid == 0) s_any = 0;
        __syncthreads();
    }
    if (tid == 0 && stats) {
        atomicAdd(&stats[0], (unsigned long long)sweeps_done);
        atomicAdd(&stats[1], 1ull);
    }

    // Ordered rotation: row k of T moves to row dst[k], into the free G buffer.
    pair_order(tid, s_row, &Gc[0][0].x, 2 * (JP + 2), s_dst);
    __syncthreads();
    for (int e = tid; e < JP * JP; e += NT) {
        const int i = e / JP, j = e % JP;
        Gn[s_dst[i]][j] = Tc[i][j];
    }
    __syncthreads();
    double2 (*const Tp)[JP + 1] = Gn;

    // ---- 3. P <- T P on MFMA. Lane l supplies T[I*16 + l%16][4kk + l/16] (read from LDS per
    // column chunk, which keeps the kernel within the registers of three workgroups per CU)
    // and P[4kk + l/16][c0 + l%16]; output element j of lane l is local row I*16 + 4j + l/16.
    {
        const int i = lane & 15, g = lane >> 4;
        int prow[8];
#pragma unroll
        for (int kk = 0; kk < 8; ++kk) prow[kk] = s_row[4 * kk + g];
        int orow[2][4];
#pragma unroll
        for (int I = 0; I < 2; ++I)
#pragma unroll
            for (int j = 0; j < 4; ++j) orow[I][j] = s_row[I * 16 + 4 * j + g];
        // 3M products: E1 = Tr Pr, E2 = Ti Pi, E3 = (Tr + Ti)(Pr + Pi); Re = E1 - E2,
        // Im = E3 - E1 - E2. Column chunks are double-buffered in registers.
        for (int pass = 0; pass < 2; ++pass) {
            double2* M = pass == 0 ? jb.X : jb.J;
            if (!M) break;
            const int cols = pass == 0 ? jb.m : jb.n;
            const long ld = pass == 0 ? jb.ldx : jb.n;
            double2 ba[8], bb[8];
#define WQC_APPLY_LOAD(C0, B)                                                        \
    {                                                                                \
        const int c = (C0) + i;                                                      \
        _Pragma("unroll") for (int kk = 0; kk < 8; ++kk)                             \
            B[kk] = (c < cols && prow[kk] >= 0) ? M[(long)prow[kk] * ld + c] : cz(); \
    }
#define WQC_APPLY_ROT(C0, B)                                                         \
    {                                                                                \
        const int c = (C0) + i;                                                      \
        double bs[8];                                                                \
        _Pragma("unroll") for (int kk = 0; kk < 8; ++kk) bs[kk] = B[kk].x + B[kk].y; \
        _Pragma("unroll") for (int I = 0; I < 2; ++I) {                              \
            d4v e1 = {0, 0, 0, 0}, e2 = e1, e3 = e1;                                 \
            _Pragma("unroll") for (int kk = 0; kk < 8; ++kk) {                       \
                const double2 t = Tp[I * 16 + i][4 * kk + g];                        \
                e1 = mfma_f64(t.x, B[kk].x, e1);                                     \
                e2 = mfma_f64(t.y, B[kk].y, e2);                                     \
                e3 = mfma_f64(t.x + t.y, bs[kk], e3);                                \
            }                                                                        \
            if (c < cols) {                                                          \
                _Pragma("unroll") for (int j = 0; j < 4; ++j) {                      \
                    const int gr = orow[I][j];                                       \
                    if (gr >= 0)                                                     \
                        M[(long)gr * ld + c] = make_double2(e1[j] - e2[j], e3[j] - e1[j] - e2[j]); \
                }                                                                    \
            }                                                                        \
        }                                                                            \
    }
            int c0 = wave * 16;
            if (c0 < cols) WQC_APPLY_LOAD(c0, ba)
            while (c0 < cols) {
                const int c1 = c0 + 16 * NW;
                if (c1 < cols) WQC_APPLY_LOAD(c1, bb)
                WQC_APPLY_ROT(c0, ba)
                if (c1 >= cols) break;
                const int c2 = c1 + 16 * NW;
                if (c2 < cols) WQC_APPLY_LOAD(c2, ba)
                WQC_APPLY_ROT(c1, bb)
                c0 = c2;
            }
#undef WQC_APPLY_LOAD
#undef WQC_APPLY_ROT
        }
    }
}

// ------------------------------------------------------------------ mixed-precision round
// Low-precision phase of the mixed split: the rows (X32) and the accumulated rotation (J32)
// live in complex64. Per block pair: the Gram on FP32 MFMA (v_mfma_f32_16x16x4f32), or with
// GB in split-BF16 on BF16 MFMA with FP32 accumulation; the eigensolve of the 32x32 Gram
// matrix in FP32; the rotation T applied to X32 and J32 on FP32 MFMA. J32 stays unitary to
// about FP32 precision; after the last low round the host lifts it to complex128 and makes it
// unitary to FP64 precision by Newton-Schulz steps, recomputes X = J X_orig in FP64, and the
// FP64 rounds finish from there.
typedef float f4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f4v mfma_f32(float a, float b, f4v c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

typedef __bf16 bf8v __attribute__((ext_vector_type(8)));

__device__ __forceinline__ f4v mfma_bf16(bf8v a, bf8v b, f4v c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// NT = 512 for rounds with at most one block pair per CU, as in k_jacobi_round. GB: the Gram
// matrix on BF16 MFMA (split-BF16, FP32 accumulation) instead of FP32 MFMA; the eigensolve and
// the rotations stay FP32.
// Rotation parameters of the low rounds in FP32 (the rotation is applied in FP32 and J is made
// unitary in FP64 once, after the last low round, so parameters accurate to FP32 suffice):
// |c| by hypotf and the tolerance against sqrt(a) sqrt(d), so nothing is squared and no entry
// of a small row underflows; hardware reciprocal and reciprocal square root (about 1 ulp).
struct Rot32 {
    int act;
    float cs, tg;
    float2 m01, m10;
};

__device__ __forceinline__ Rot32 rot_params32(float a, float d, float2 c, float tol, float low) {
    Rot32 r;
    const float g = hypotf(c.x, c.y);
    const bool act = g > 0.f && fminf(a, d) > low && g > tol * (sqrtf(a) * sqrtf(d));
    const float gs = act ? g : 1.f;
    const float rg = __builtin_amdgcn_rcpf(gs);
    const float zeta = 0.5f * (d - a) * rg;
    const float w = fmaf(zeta, zeta, 1.f);
    const float t = copysignf(__builtin_amdgcn_rcpf(fabsf(zeta) + w * __builtin_amdgcn_rsqf(w)), zeta);
    const float cs = __builtin_amdgcn_rsqf(fmaf(t, t, 1.f));
    const float sn = cs * t;
    const float phx = c.x * rg, phy = c.y * rg;
    r.act = act ? 1 : 0;
    r.cs = act ? cs : 1.f;
    r.m01 = act ? make_float2(-phx * sn, -phy * sn) : make_float2(0.f, 0.f);
    r.m10 = act ? make_float2(phx * sn, -phy * sn) : make_float2(0.f, 0.f);
    r.tg = act ? t * gs : 0.f;
    return r;
}

// As rot_from_lane: the row-block rotation from the lane that derived it; tg stays the lane's
// own (used only on a diagonal tile, whose row rotation is the lane's own).
__device__ __forceinline__ Rot32 rot_from_lane32(const Rot32& own, int src) {
    Rot32 r;
    r.act = __shfl(own.act, src, 64);
    r.cs = __shfl(own.cs, src, 64);
    r.tg = own.tg;
    r.m01.x = __shfl(own.m01.x, src, 64);
    r.m01.y = __shfl(own.m01.y, src, 64);
    r.m10 = make_float2(-r.m01.x, r.m01.y);
    return r;
}

template <int NT, bool GB>
__global__ __launch_bounds__(NT, NT == 256 ? 3 : 1) void k_jacobi_round_lo(const JacobiLoJob* __restrict__ jobs,
                                                         int round, float tol_base, double nu2_base,
                                                         float rel_floor, int inner_cap,
                                                         int* __restrict__ rotated,
                                                         const int* __restrict__ done,
                                                         unsigned long long* __restrict__ stats,
                                                         unsigned long long* __restrict__ offmax) {
    constexpr int NW = NT / 64;
    const int job = blockIdx.y;
    if (done[job]) return;
    const JacobiLoJob jb = jobs[job];
    const float tol = (float)((double)tol_base * sqrt((double)jb.m));
    const double nu2 = nu2_base / (double)jb.n;
    const int nb = (jb.n + JB - 1) / JB;
    const int nbe = nb + (nb & 1);
    if (round >= nbe - 1) return;
    const int kpair = blockIdx.x;
    if (kpair >= nbe / 2) return;
    int bi, bj;
    if (kpair == 0) {
        bi = nbe - 1;
        bj = round;
    } else {
        bi = (round + kpair) % (nbe - 1);
        bj = (round - kpair + nbe - 1) % (nbe - 1);
    }

    __shared__ float2 G[JP][JP + 1];
    __shared__ float2 T[JP][JP + 1];
    // Partial Gram tiles of two waves (four with eight waves); afterwards the second G buffer of
    // the eigensolve.
    __shared__ __align__(16) float s_aux[NW == 8 ? 4 * 6 * 4 * 64 : 2 * 6 * 4 * 64];
    __shared__ int s_row[JP];
    __shared__ int s_any;
    __shared__ int s_intra;
    __shared__ float s_low;
    __shared__ int s_dst[JP];
    float (*Gp)[6][4][64] = reinterpret_cast<float (*)[6][4][64]>(s_aux);
    float2 (*G2)[JP + 1] = reinterpret_cast<float2 (*)[JP + 1]>(s_aux);

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid < JP) {
        const int b = tid < JB ? bi : bj;
        const int r = b * JB + (tid & (JB - 1));
        s_row[tid] = (b < nb && r < jb.n) ? r : -1;
    }
    if (tid == 0) {
        s_any = 0;
        s_intra = 0;
    }
    __syncthreads();

    {
        const int i = lane & 15, g = lane >> 4;
        const int r0 = s_row[i], r1 = s_row[16 + i];
        const float2* x0 = r0 >= 0 ? jb.X + (long)r0 * jb.ldx : nullptr;
        const float2* x1 = r1 >= 0 ? jb.X + (long)r1 * jb.ldx : nullptr;
        // 3M products and register double-buffered column chunks, as in the FP64 round.
        f4v p00 = {0, 0, 0, 0}, q00 = p00, t00 = p00, p01 = p00, q01 = p00, t01 = p00, p11 = p00, q11 = p00, t11 = p00;
        const float2 zf = make_float2(0.f, 0.f);
        if constexpr (GB) {
            // Split-BF16 Gram: every FP32 operand a = a_hi + a_lo (two bf16, round to nearest)
            // and every real product a b ~ a_hi b_hi + a_hi b_lo + a_lo b_hi on
            // v_mfma_f32_16x16x32_bf16 with FP32 accumulation (about 2^-17 relative per
            // product). Lane (i, g) supplies columns c0 + 8g .. c0 + 8g + 7 of local rows i and
            // 16 + i as both operands, so the k order inside the instruction is the same for A
            // and B; 32 columns per wave and chunk.
            for (int c0 = wave * 32; c0 < jb.m; c0 += 32 * NW) {
                bf8v uh[4], ul[4], vh[4], vl[4];  // re, im, re + im, re - im
#pragma unroll
                for (int s = 0; s < 8; ++s) {
                    const int c = c0 + 8 * g + s;
                    const float2 u = (x0 && c < jb.m) ? x0[c] : zf;
                    const float2 v = (x1 && c < jb.m) ? x1[c] : zf;
                    const float uv[4] = {u.x, u.y, u.x + u.y, u.x - u.y};
                    const float vv[4] = {v.x, v.y, v.x + v.y, v.x - v.y};
#pragma unroll
                    for (int t = 0; t < 4; ++t) {
                        const __bf16 a = (__bf16)uv[t];
                        uh[t][s] = a;
                        ul[t][s] = (__bf16)(uv[t] - (float)a);
                        const __bf16 b = (__bf16)vv[t];
                        vh[t][s] = b;
                        vl[t][s] = (__bf16)(vv[t] - (float)b);
                    }
                }
#define WQC_MFMA_X3(ACC, AH, AL, BH, BL) \
    ACC = mfma_bf16(AL, BH, ACC);        \
    ACC = mfma_bf16(AH, BL, ACC);        \
    ACC = mfma_bf16(AH, BH, ACC);
                WQC_MFMA_X3(p00, uh[0], ul[0], uh[0], ul[0])
                WQC_MFMA_X3(q00, uh[1], ul[1], uh[1], ul[1])
                WQC_MFMA_X3(t00, uh[2], ul[2], uh[3], ul[3])
                WQC_MFMA_X3(p01, uh[0], ul[0], vh[0], vl[0])
                WQC_MFMA_X3(q01, uh[1], ul[1], vh[1], vl[1])
                WQC_MFMA_X3(t01, uh[2], ul[2], vh[3], vl[3])
                WQC_MFMA_X3(p11, vh[0], vl[0], vh[0], vl[0])
                WQC_MFMA_X3(q11, vh[1], vl[1], vh[1], vl[1])
                WQC_MFMA_X3(t11, vh[2], vl[2], vh[3], vl[3])
#undef WQC_MFMA_X3
            }
        } else {
        float2 ua[4], va[4], ub[4], vb[4];
#define WQC_GRAM_LOAD(C0, U, V)                                  \
    _Pragma("unroll") for (int s = 0; s < 4; ++s) {              \
        const int c = (C0) + 4 * g + s;                          \
        U[s] = (x0 && c < jb.m) ? x0[c] : zf;                    \
        V[s] = (x1 && c < jb.m) ? x1[c] : zf;                    \
    }
#define WQC_GRAM_ACC(U, V)                                       \
    _Pragma("unroll") for (int s = 0; s < 4; ++s) {              \
        const float up = U[s].x + U[s].y, um = U[s].x - U[s].y;  \
        const float vp = V[s].x + V[s].y, vm = V[s].x - V[s].y;  \
        p00 = mfma_f32(U[s].x, U[s].x, p00);                     \
        q00 = mfma_f32(U[s].y, U[s].y, q00);                     \
        t00 = mfma_f32(up, um, t00);                             \
        p01 = mfma_f32(U[s].x, V[s].x, p01);                     \
        q01 = mfma_f32(U[s].y, V[s].y, q01);                     \
        t01 = mfma_f32(up, vm, t01);                             \
        p11 = mfma_f32(V[s].x, V[s].x, p11);                     \
        q11 = mfma_f32(V[s].y, V[s].y, q11);                     \
        t11 = mfma_f32(vp, vm, t11);                             \
    }
        int c0 = wave * 16;
        if (c0 < jb.m) WQC_GRAM_LOAD(c0, ua, va)
        while (c0 < jb.m) {
            const int c1 = c0 + 16 * NW;
            if (c1 < jb.m) WQC_GRAM_LOAD(c1, ub, vb)
            WQC_GRAM_ACC(ua, va)
            if (c1 >= jb.m) break;
            const int c2 = c1 + 16 * NW;
            if (c2 < jb.m) WQC_GRAM_LOAD(c2, ua, va)
            WQC_GRAM_ACC(ub, vb)
            c0 = c2;
        }
#undef WQC_GRAM_LOAD
#undef WQC_GRAM_ACC
        }
        f4v a00r = p00 + q00, a00i = t00 - p00 + q00;
        f4v a01r = p01 + q01, a01i = t01 - p01 + q01;
        f4v a11r = p11 + q11, a11i = t11 - p11 + q11;
#define WQC_GP_ALL(OP) OP(0, a00r) OP(1, a00i) OP(2, a01r) OP(3, a01i) OP(4, a11r) OP(5, a11i)
#define WQC_GP_STORE(t, v) for (int j = 0; j < 4; ++j) Gp[slot][t][j][lane] = v[j];
#define WQC_GP_ADD(t, v) for (int j = 0; j < 4; ++j) v[j] += Gp[slot][t][j][lane];
        if constexpr (NW == 8) {
            // (w0 + w4), (w1 + w5), (w2 + w6), (w3 + w7) first.
            if (wave >= 4) {
                const int slot = wave - 4;
                WQC_GP_ALL(WQC_GP_STORE)
            }
            __syncthreads();
            if (wave < 4) {
                const int slot = wave;
                WQC_GP_ALL(WQC_GP_ADD)
            }
            __syncthreads();
        }
        if (wave >= 2 && wave < 4) {
            const int slot = wave - 2;
            WQC_GP_ALL(WQC_GP_STORE)
        }
        __syncthreads();
        if (wave < 2) {
            const int slot = wave;
            WQC_GP_ALL(WQC_GP_ADD)
        }
        __syncthreads();
        if (wave == 1) {
            const int slot = 0;
            WQC_GP_ALL(WQC_GP_STORE)
        }
        __syncthreads();
#undef WQC_GP_ALL
#undef WQC_GP_STORE
#undef WQC_GP_ADD
        if (wave == 0) {
            // MFMA f32 output: element j of lane l is row 4*(l/16) + j, column l%16.
            const int c = lane & 15;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int rr = 4 * (lane >> 4) + j;
                const float g00r = a00r[j] + Gp[0][0][j][lane], g00i = a00i[j] + Gp[0][1][j][lane];
                const float g01r = a01r[j] + Gp[0][2][j][lane], g01i = a01i[j] + Gp[0][3][j][lane];
                const float g11r = a11r[j] + Gp[0][4][j][lane], g11i = a11i[j] + Gp[0][5][j][lane];
                G[rr][c] = make_float2(g00r, g00i);
                G[rr][16 + c] = make_float2(g01r, g01i);
                G[16 + c][rr] = make_float2(g01r, -g01i);
                G[16 + rr][16 + c] = make_float2(g11r, g11i);
            }
        }
    }
    __syncthreads();
    // Rows whose squared norm is below rel_floor of the pair's largest are left to the FP64
    // finish: FP32 rotations add noise of about 1e-7 of the largest row to them, so they could
    // not reach the relative tolerance here.
    if (wave == 0) {
        float mx = lane < JP ? G[lane][lane].x : 0.f;
#pragma unroll
        for (int o = JP / 2; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o, 64));
        if (lane == 0) s_low = fmaxf((float)nu2, mx * rel_floor);
    }
    __syncthreads();
    const float low = s_low;
    {
        bool off = false, intra = false;
        double worst = 0.0;
        for (int e = tid; e < JP * JP; e += NT) {
            const int i = e / JP, j = e % JP;
            if (i < j) {
                const float2 c = G[i][j];
                const float gg = hypotf(c.x, c.y);
                const float di = G[i][i].x, dj = G[j][j].x;
                if (gg > 0.f && fminf(di, dj) > low) {
                    const float den = sqrtf(fmaxf(di, 0.f) * fmaxf(dj, 0.f));
                    if (gg > tol * den) {
                        off = true;
                        if ((i < JB) == (j < JB)) intra = true;
                    }
                    if (offmax) worst = fmax(worst, (double)(gg / den));
                }
            }
        }
        if (off) s_any = 1;
        if (intra) s_intra = 1;
        if (offmax && worst > 0.0) atomicMax(&offmax[job], (unsigned long long)__double_as_longlong(worst));
    }
    __syncthreads();
    const int need = s_any;
    const int intra0 = s_intra;
    __syncthreads();
    if (!need) return;
    if (tid == 0) {
        rotated[job] = 1;
        s_any = 0;
    }
    for (int e = tid; e < JP * JP; e += NT) {
        const int i = e / JP, j = e % JP;
        T[i][j] = make_float2(i == j ? 1.f : 0.f, 0.f);
    }
    __syncthreads();

    float2 (*Gc)[JP + 1] = G;
    float2 (*const Tc)[JP + 1] = T;  // updated in place: a step's tiles are disjoint
    float2 (*Gn)[JP + 1] = G2;
    int sweeps_done = 0;
    const int bk = tid >> 4, bk2 = tid & 15;
    auto fm = [](float2 a, float2 b) { return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); };
    auto fa = [](float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); };
    auto fs = [](float2 a, float k) { return make_float2(a.x * k, a.y * k); };
    auto fc = [](float2 a) { return make_float2(a.x, -a.y); };
    for (int sweep = 0; sweep < inner_cap; ++sweep) {
        ++sweeps_done;
        // The pair's two blocks come out of their previous rounds mutually orthogonal inside;
        // when the check found no block-internal pair above tolerance, the first sweep
        // rotates only the JB x JB cross pairs (JB steps instead of 2 JB - 1).
        const bool cross = sweep == 0 && !intra0;
        const int nsteps = cross ? JB : JP - 1;
        for (int step = 0; step < nsteps; ++step) {
            if (tid < 256) {  // the eigensolve's 16 x 16 tiles of 2 x 2
                int p1, q1, p2, q2;
                if (cross) {
                    cross_pair(step, bk, p1, q1);
                    cross_pair(step, bk2, p2, q2);
                } else {
                    circle_pair(step, bk, p1, q1);
                    circle_pair(step, bk2, p2, q2);
                }
                const Rot32 r2 = rot_params32(Gc[p2][p2].x, Gc[q2][q2].x, Gc[p2][q2], tol, low);
                const Rot32 r1 = rot_from_lane32(r2, bk & 15);
                float2 h00 = Gc[p1][p2], h01 = Gc[p1][q2], h10 = Gc[q1][p2], h11 = Gc[q1][q2];
                if (r1.act) {
                    const float cs = r1.cs;
                    const float2 m01 = r1.m01, m10 = r1.m10;
                    const float2 x00 = fa(fs(h00, cs), fm(m01, h10)), x01 = fa(fs(h01, cs), fm(m01, h11));
                    const float2 x10 = fa(fm(m10, h00), fs(h10, cs)), x11 = fa(fm(m10, h01), fs(h11, cs));
                    h00 = x00;
                    h01 = x01;
                    h10 = x10;
                    h11 = x11;
                    // T changes only with the row rotation.
                    const float2 t00 = Tc[p1][p2], t01 = Tc[p1][q2], t10 = Tc[q1][p2], t11 = Tc[q1][q2];
                    Tc[p1][p2] = fa(fs(t00, cs), fm(m01, t10));
                    Tc[p1][q2] = fa(fs(t01, cs), fm(m01, t11));
                    Tc[q1][p2] = fa(fm(m10, t00), fs(t10, cs));
                    Tc[q1][q2] = fa(fm(m10, t01), fs(t11, cs));
                }
                if (r2.act) {
                    const float cs = r2.cs;
                    const float2 c01 = fc(r2.m01), c10 = fc(r2.m10);
                    const float2 n00 = fa(fs(h00, cs), fm(h01, c01)), n01 = fa(fm(h00, c10), fs(h01, cs));
                    const float2 n10 = fa(fs(h10, cs), fm(h11, c01)), n11 = fa(fm(h10, c10), fs(h11, cs));
                    h00 = n00;
                    h01 = n01;
                    h10 = n10;
                    h11 = n11;
                }
                if (bk == bk2 && r1.act) {
                    h00 = make_float2(Gc[p1][p1].x - r1.tg, 0.f);
                    h11 = make_float2(Gc[q1][q1].x + r1.tg, 0.f);
                    h01 = make_float2(0.f, 0.f);
                    h10 = h01;
                    s_any = 1;
                }
                Gn[p1][p2] = h00;
                Gn[p1][q2] = h01;
                Gn[q1][p2] = h10;
                Gn[q1][q2] = h11;
            }
            __syncthreads();
            float2 (*sw)[JP + 1] = Gc;
            Gc = Gn;
            Gn = sw;
        }
        const int again = s_any;
        __syncthreads();
        if (!again) break;
        if (tid == 0) s_any = 0;
        __syncthreads();
    }
    if (tid == 0 && stats) {
        atomicAdd(&stats[0], (unsigned long long)sweeps_done);
        atomicAdd(&stats[1], 1ull);
    }
    // Ordered rotation (as in the FP64 round): row k of T moves to row dst[k], into the free G
    // buffer.
    pair_order(tid, s_row, &Gc[0][0].x, 2 * (JP + 2), s_dst);
    __syncthreads();
    for (int e = tid; e < JP * JP; e += NT) {
        const int i = e / JP, j = e % JP;
        Gn[s_dst[i]][j] = Tc[i][j];
    }
    __syncthreads();
    float2 (*const Tp)[JP + 1] = Gn;

    // X32 <- T X32 and J32 <- T J32 on FP32 MFMA with 3M products (E1 = Tr Pr, E2 = Ti Pi,
    // E3 = (Tr + Ti)(Pr + Pi); Re = E1 - E2, Im = E3 - E1 - E2). J stays unitary to FP32
    // precision; the host makes it unitary in FP64 once, after the last low round. T is read
    // from LDS per column chunk (three workgroups per CU fit in the registers).
    {
        const int i = lane & 15, g = lane >> 4;
        int prow[8];
#pragma unroll
        for (int kk = 0; kk < 8; ++kk) prow[kk] = s_row[4 * kk + g];
        // MFMA f32 output: element j of lane l is row 4 (l / 16) + j, column l % 16.
        int orow[2][4];
#pragma unroll
        for (int I = 0; I < 2; ++I)
#pragma unroll
            for (int j = 0; j < 4; ++j) orow[I][j] = s_row[I * 16 + 4 * g + j];
        for (int pass = 0; pass < 2; ++pass) {
            float2* M = pass == 0 ? jb.X : jb.J;
            const int cols = pass == 0 ? jb.m : jb.n;
            const long ld = pass == 0 ? jb.ldx : jb.n;
#define WQC_APPLY_LOAD(C0, B)                                                                      \
    {                                                                                              \
        const int c = (C0) + i;                                                                    \
        _Pragma("unroll") for (int kk = 0; kk < 8; ++kk)                                           \
            B[kk] = (c < cols && prow[kk] >= 0) ? M[(long)prow[kk] * ld + c] : make_float2(0.f, 0.f); \
    }
#define WQC_APPLY_ROT(C0, B)                                                                       \
    {                                                                                              \
        const int c = (C0) + i;                                                                    \
        float bs[8];                                                                               \
        _Pragma("unroll") for (int kk = 0; kk < 8; ++kk) bs[kk] = B[kk].x + B[kk].y;               \
        _Pragma("unroll") for (int I = 0; I < 2; ++I) {                                            \
            f4v e1 = {0, 0, 0, 0}, e2 = e1, e3 = e1;                                               \
            _Pragma("unroll") for (int kk = 0; kk < 8; ++kk) {                                     \
                const float2 t = Tp[I * 16 + i][4 * kk + g];                                       \
                e1 = mfma_f32(t.x, B[kk].x, e1);                                                   \
                e2 = mfma_f32(t.y, B[kk].y, e2);                                                   \
                e3 = mfma_f32(t.x + t.y, bs[kk], e3);                                              \
            }                                                                                      \
            if (c < cols) {                                                                        \
                _Pragma("unroll") for (int j = 0; j < 4; ++j) {                                    \
                    const int gr = orow[I][j];                                                     \
                    if (gr >= 0) M[(long)gr * ld + c] = make_float2(e1[j] - e2[j], e3[j] - e1[j] - e2[j]); \
                }                                                                                  \
            }                                                                                      \
        }                                                                                          \
    }
            float2 ba[8], bb[8];
            int c0 = wave * 16;
            if (c0 < cols) WQC_APPLY_LOAD(c0, ba)
            while (c0 < cols) {
                const int c1 = c0 + 16 * NW;
                if (c1 < cols) WQC_APPLY_LOAD(c1, bb)
                WQC_APPLY_ROT(c0, ba)
                if (c1 >= cols) break;
                const int c2 = c1 + 16 * NW;
                if (c2 < cols) WQC_APPLY_LOAD(c2, ba)
                WQC_APPLY_ROT(c1, bb)
                c0 = c2;
            }
#undef WQC_APPLY_LOAD
#undef WQC_APPLY_ROT
        }
    }
}

// X32 = (complex64) X
__global__ __launch_bounds__(256) void k_to_f32(const double2* const* __restrict__ src,
                                                float2* const* __restrict__ dst,
                                                const long* __restrict__ counts) {
    const int job = blockIdx.y;
    const double2* s = src[job];
    float2* d = dst[job];
    const long n = counts[job];
    for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < n; e += (long)gridDim.x * 256) {
        const double2 v = s[e];
        d[e] = make_float2((float)v.x, (float)v.y);
    }
}

// (complex128) of a complex64 buffer.
__global__ __launch_bounds__(256) void k_to_f64(const float2* const* __restrict__ src,
                                                double2* const* __restrict__ dst,
                                                const long* __restrict__ counts) {
    const int job = blockIdx.y;
    const float2* s = src[job];
    double2* d = dst[job];
    const long n = counts[job];
    for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < n; e += (long)gridDim.x * 256) {
        const float2 v = s[e];
        d[e] = make_double2((double)v.x, (double)v.y);
    }
}

// jobs[j].dst[e] = jobs[j].src[e], e < jobs[j].n (doubles): the per-job device copies of a batch
// (saved rows, J copies, sigma gathers) as one launch instead of one copy per job.
__global__ __launch_bounds__(256) void k_copy_batch(const CopyJob* __restrict__ jobs) {
    const CopyJob jb = jobs[blockIdx.y];
    for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < jb.n; e += (long)gridDim.x * 256) jb.dst[e] = jb.src[e];
}

// Probe residuals of the split check: out[p] = sum_i |d[i, p] - r[i, p]|^2 for the two probe
// columns p of the n x 2 row-major d and r, one workgroup per job, fixed summation order.
__global__ __launch_bounds__(256) void k_probe_residual(const ProbeJob* __restrict__ jobs) {
    const ProbeJob jb = jobs[blockIdx.x];
    double acc[2] = {0.0, 0.0};
    for (int i = threadIdx.x; i < jb.n; i += 256)
#pragma unroll
        for (int p = 0; p < 2; ++p) {
            const double2 e = csub(jb.d[2 * i + p], jb.r[2 * i + p]);
            acc[p] = fma(e.x, e.x, acc[p]);
            acc[p] = fma(e.y, e.y, acc[p]);
        }
    __shared__ double ws[2][4];
#pragma unroll
    for (int p = 0; p < 2; ++p) {
        const double v = wave_sum(acc[p]);
        if ((threadIdx.x & 63) == 0) ws[p][threadIdx.x >> 6] = v;
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        const int p = threadIdx.x;
        jb.out[p] = ((ws[p][0] + ws[p][1]) + ws[p][2]) + ws[p][3];
    }
}

// sigma[row] = ||X[row, :]||, one wave per row.
__global__ __launch_bounds__(256) void k_row_norms(const JacobiJob* __restrict__ jobs,
                                                   double* __restrict__ sigma,
                                                   const long* __restrict__ off) {
    const int job = blockIdx.y;
    const JacobiJob jb = jobs[job];
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= jb.n) return;
    const int lane = threadIdx.x & 63;
    double acc = 0.0;
    const double2* x = jb.X + (long)row * jb.ldx;
    for (int c = lane; c < jb.m; c += 64) {
        const double2 v = x[c];
        acc = fma(v.x, v.x, acc);
        acc = fma(v.y, v.y, acc);
    }
    acc = wave_sum(acc);
    if (lane == 0) sigma[off[job] + row] = sqrt(acc);
}

// ------------------------------------------------------------------ batched ZGEMM
// ------------------------------------------------------------------ MFMA GEMM
// Batched complex FP64 GEMM C = alpha op(A) op(B) + beta C on v_mfma_f64_16x16x4f64 with 3M
// products (E1 = Ar Br, E2 = Ai Bi, E3 = (Ar + Ai)(Br + Bi); Re = E1 - E2, Im = E3 - E1 - E2).
// A workgroup computes a TL x TL tile (TL = 64: each wave a 32 x 32 quarter of 2 x 2 MFMA tiles;
// TL = 32: each wave one 16 x 16 MFMA tile, four times the workgroups for skinny products); K
// advances in steps of 16 staged through LDS as real / imaginary planes (op C transposes and
// conjugates while staging). Edges are zero-padded. Every output element sees the same sequence
// of MFMA accumulations over k at either tile size, so the two give bit-identical results.
constexpr int MG = 64, MGK = 16;

template <bool PARTIAL, int TL>
__global__ __launch_bounds__(256, 2) void k_gemm_mfma(const GemmJob* __restrict__ jobs, int splits) {
    const GemmJob jb = jobs[PARTIAL ? blockIdx.z / splits : blockIdx.z];
    const int m0 = blockIdx.y * TL, n0 = blockIdx.x * TL;
    if (m0 >= jb.M || n0 >= jb.N) return;
    int kbeg = 0, kend = jb.K;
    if (PARTIAL) {
        const int slice = blockIdx.z % splits;
        const int ks = (((jb.K + splits - 1) / splits) + MGK - 1) / MGK * MGK;
        kbeg = slice * ks;
        kend = min(jb.K, kbeg + ks);
    }
    __shared__ double Ar[TL][MGK + 1], Ai[TL][MGK + 1], Br[MGK][TL + 1], Bi[MGK][TL + 1];
    constexpr int WT = TL / 2, NA = TL / 32;  // each wave: a WT x WT quarter of NA x NA MFMA tiles
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave >> 1) * WT, wn = (wave & 1) * WT;
    const int li = lane & 15, lk = lane >> 4;
    d4v e1[NA][NA], e2[NA][NA], e3[NA][NA];
#pragma unroll
    for (int a = 0; a < NA; ++a)
#pragma unroll
        for (int b = 0; b < NA; ++b) {
            e1[a][b] = d4v{0, 0, 0, 0};
            e2[a][b] = e1[a][b];
            e3[a][b] = e1[a][b];
        }
    // The next K tile is fetched into registers while the MFMAs consume the current one from
    // LDS (the accumulation order over k is unchanged).
    constexpr int PER = (TL * MGK) / 256;
    double2 ra[PER], rb[PER];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int e = tid + 256 * u;
            if (jb.opA == 0) {
                const int m = m0 + e / MGK, k = k0 + e % MGK;
                ra[u] = (m < jb.M && k < kend) ? jb.A[(long)m * jb.lda + k] : cz();
            } else {
                const int m = m0 + e % TL, k = k0 + e / TL;
                ra[u] = (m < jb.M && k < kend) ? cconj(jb.A[(long)k * jb.lda + m]) : cz();
            }
            if (jb.opB == 0) {
                const int n = n0 + e % TL, k = k0 + e / TL;
                rb[u] = (n < jb.N && k < kend) ? jb.B[(long)k * jb.ldb + n] : cz();
            } else {
                const int n = n0 + e / MGK, k = k0 + e % MGK;
                rb[u] = (n < jb.N && k < kend) ? cconj(jb.B[(long)n * jb.ldb + k]) : cz();
            }
        }
    };
    if (kbeg < kend) fetch(kbeg);
    for (int k0 = kbeg; k0 < kend; k0 += MGK) {
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int e = tid + 256 * u;
            const int am = jb.opA == 0 ? e / MGK : e % TL, ak = jb.opA == 0 ? e % MGK : e / TL;
            Ar[am][ak] = ra[u].x;
            Ai[am][ak] = ra[u].y;
            const int bn = jb.opB == 0 ? e % TL : e / MGK, bk = jb.opB == 0 ? e / TL : e % MGK;
            Br[bk][bn] = rb[u].x;
            Bi[bk][bn] = rb[u].y;
        }
        __syncthreads();
        if (k0 + MGK < kend) fetch(k0 + MGK);
#pragma unroll
        for (int ks = 0; ks < MGK / 4; ++ks) {
            const int kk = 4 * ks + lk;
            double ar[NA], ai[NA], as[NA], br[NA], bi[NA], bs[NA];
#pragma unroll
            for (int a = 0; a < NA; ++a) {
                ar[a] = Ar[wm + 16 * a + li][kk];
                ai[a] = Ai[wm + 16 * a + li][kk];
                as[a] = ar[a] + ai[a];
                br[a] = Br[kk][wn + 16 * a + li];
                bi[a] = Bi[kk][wn + 16 * a + li];
                bs[a] = br[a] + bi[a];
            }
#pragma unroll
            for (int a = 0; a < NA; ++a)
#pragma unroll
                for (int b = 0; b < NA; ++b) {
                    e1[a][b] = mfma_f64(ar[a], br[b], e1[a][b]);
                    e2[a][b] = mfma_f64(ai[a], bi[b], e2[a][b]);
                    e3[a][b] = mfma_f64(as[a], bs[b], e3[a][b]);
                }
        }
        __syncthreads();
    }
    // MFMA f64 output: element j of lane l is row 4j + l/16, column l%16 of the 16 x 16 tile.
#pragma unroll
    for (int a = 0; a < NA; ++a)
#pragma unroll
        for (int b = 0; b < NA; ++b)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int m = m0 + wm + 16 * a + 4 * j + lk, n = n0 + wn + 16 * b + li;
                if (m >= jb.M || n >= jb.N) continue;
                const double re = e1[a][b][j] - e2[a][b][j], im = e3[a][b][j] - e1[a][b][j] - e2[a][b][j];
                if (PARTIAL) {
                    jb.part[((long)(blockIdx.z % splits) * jb.M + m) * jb.N + n] = make_double2(re, im);
                    continue;
                }
                double2* c = jb.C + (long)m * jb.ldc + n;
                double2 v = cmul(jb.alpha, make_double2(re, im));
                if (jb.beta.x != 0.0 || jb.beta.y != 0.0) cmac(v, jb.beta, *c);
                *c = v;
            }
}

// Sums the split-K partials in slice order, then C = alpha * sum + beta * C.
__global__ __launch_bounds__(256) void k_gemm_splitk_reduce(const GemmJob* __restrict__ jobs, int splits) {
    const GemmJob jb = jobs[blockIdx.y];
    const long total = (long)jb.M * jb.N;
    for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
        double2 acc = cz();
        for (int sl = 0; sl < splits; ++sl) acc = cadd(acc, jb.part[(long)sl * total + e]);
        const int m = (int)(e / jb.N), n = (int)(e - (long)m * jb.N);
        double2* c = jb.C + (long)m * jb.ldc + n;
        double2 v = cmul(jb.alpha, acc);
        if (jb.beta.x != 0.0 || jb.beta.y != 0.0) cmac(v, jb.beta, *c);
        *c = v;
    }
}

// ------------------------------------------------------------------ factors
__global__ __launch_bounds__(256) void k_factor(const FactorJob* __restrict__ jobs) {
    const FactorJob jb = jobs[blockIdx.y];
    const bool left = blockIdx.z == 0;
    // modes 2 and 3 write W (ell x chi) into `left`; its row count is the Jacobi dimension.
    const long nrows = left ? (jb.mode == 2 ? (long)jb.ldj : jb.mode == 3 ? (long)jb.ldx : (long)jb.rows)
                            : (long)jb.chi;
    const bool swapped = jb.mode == 1 || jb.mode == 3;
    const long ncols = left ? (long)jb.chi : (long)jb.cols;
    const long total = nrows * ncols;
    for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
        const long i = e / ncols, j = e - i * ncols;
        double2 v = cz();
        if (left) {
            const int p = jb.perm[j];
            const double s = jb.sigma[p];
            if (s > 0.0) {
                if (swapped)
                    v = cscale(cconj(jb.X[(long)p * jb.ldx + i]), 1.0 / sqrt(s));
                else
                    v = cscale(cconj(jb.J[(long)p * jb.ldj + i]), sqrt(s));
            }
            jb.left[e] = v;
        } else {
            const int p = jb.perm[i];
            const double s = jb.sigma[p];
            if (s > 0.0) {
                if (swapped)
                    v = cscale(jb.J[(long)p * jb.ldj + j], sqrt(s));
                else
                    v = cscale(jb.X[(long)p * jb.ldx + j], 1.0 / sqrt(s));
            }
            jb.right[e] = v;
        }
    }
}

// ------------------------------------------------------------------ Householder LQ
// One workgroup (8 waves) per job factors rows j0 .. j0+nbp-1 of X (see LqJob). Thread tid
// holds columns b0 + tid + LQ_NT k (k < CPT) of the panel rows in registers (b0 = the
// block's first row, so the block's Ubc rows are written from column b0 on). For panel row
// t (diagonal column j = j0 + t), zlarfg on the column vector conj(x[t, j:]) gives beta
// (real), tau and u (u_j = 1) with x[t, j:] (I - tau u u^H) = beta e_0; the panel rows below
// take the same reflector, and T gains column t, T[0:t, t] = -tau T[0:t, 0:t] U^H u. Once row
// t is done its slots right of the diagonal hold u. Sums are fixed-order (wave shuffles,
// then the wave partials in order), so the factorization is bit-reproducible.
template <int V>
__device__ __forceinline__ void lq_block_sum(double (&v)[V], double (*red)[V], double* fin, int tid) {
    const int lane = tid & 63, wave = tid >> 6;
#pragma unroll
    for (int i = 0; i < V; ++i) v[i] = wave_sum(v[i]);
    if (lane == 0) {
#pragma unroll
        for (int i = 0; i < V; ++i) red[wave][i] = v[i];
    }
    __syncthreads();
    if (tid < V) {
        double acc = 0.0;
        for (int w = 0; w < LQ_NT / 64; ++w) acc += red[w][tid];
        fin[tid] = acc;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < V; ++i) v[i] = fin[i];
}

// Block sum of 2 * LQ_NB = 32 values per thread by a reduce-scatter inside each wave (stage k
// exchanges half of the remaining values with lane ^ (32 >> k): 31 value shuffles instead of
// 32 full butterflies), then the wave partials summed in wave order. Every thread returns
// all 32 sums. Fixed order, so bit-reproducible.
__device__ __forceinline__ void lq_block_sum32(double (&v)[2 * LQ_NB], double (*red)[2 * LQ_NB], double* fin,
                                               int tid) {
    static_assert(2 * LQ_NB == 32, "reduce-scatter is written for 32 values");
    const int lane = tid & 63, wave = tid >> 6;
    double a[16], b[8], c[4], d[2];
    const bool h5 = lane & 32, h4 = lane & 16, h3 = lane & 8, h2 = lane & 4, h1 = lane & 2;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const double send = h5 ? v[i] : v[16 + i];
        const double keep = h5 ? v[16 + i] : v[i];
        a[i] = keep + __shfl_xor(send, 32, 64);
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const double send = h4 ? a[i] : a[8 + i];
        const double keep = h4 ? a[8 + i] : a[i];
        b[i] = keep + __shfl_xor(send, 16, 64);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const double send = h3 ? b[i] : b[4 + i];
        const double keep = h3 ? b[4 + i] : b[i];
        c[i] = keep + __shfl_xor(send, 8, 64);
    }
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const double send = h2 ? c[i] : c[2 + i];
        const double keep = h2 ? c[2 + i] : c[i];
        d[i] = keep + __shfl_xor(send, 4, 64);
    }
    double e;
    {
        const double send = h1 ? d[0] : d[1];
        const double keep = h1 ? d[1] : d[0];
        e = keep + __shfl_xor(send, 2, 64);
    }
    e += __shfl_xor(e, 1, 64);
    // Lane bits 5..1 select the value: 16 h5 + 8 h4 + 4 h3 + 2 h2 + h1.
    const int idx = (h5 ? 16 : 0) + (h4 ? 8 : 0) + (h3 ? 4 : 0) + (h2 ? 2 : 0) + (h1 ? 1 : 0);
    if ((lane & 1) == 0) red[wave][idx] = e;
    __syncthreads();
    if (tid < 2 * LQ_NB) {
        double acc = 0.0;
        for (int w = 0; w < LQ_NT / 64; ++w) acc += red[w][tid];
        fin[tid] = acc;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 2 * LQ_NB; ++i) v[i] = fin[i];
}

// Shared state of one panel factorization.
struct LqShared {
    double red1[LQ_NT / 64][1];
    double fin1[1];
    double redv[LQ_NT / 64][2 * LQ_NB];
    double finv[2 * LQ_NB];
    double2 alpha;
    double2 T[LQ_NB][LQ_NB];
    double beta[LQ_NB];
};

// Step T of the panel (row T, diagonal column j0 + T); a template so that every register
// row index is a compile-time constant.
template <int CPT, int T>
__device__ __forceinline__ void lq_step(double2 (&x)[LQ_NB][CPT], LqShared& sh, int b0, int j0, int m, int nbp, int tid) {
    if constexpr (T < LQ_NB) {
        if (T < nbp) {
            const double2 one = make_double2(1.0, 0.0);
            const int j = j0 + T;
            double nrm[1] = {0.0};
#pragma unroll
            for (int k = 0; k < CPT; ++k) {
                const int c = b0 + tid + k * LQ_NT;
                if (c > j && c < m) nrm[0] = fma(x[T][k].x, x[T][k].x, fma(x[T][k].y, x[T][k].y, nrm[0]));
                if (c == j) sh.alpha = x[T][k];
            }
            lq_block_sum<1>(nrm, sh.red1, sh.fin1, tid);
            const double2 ac = cconj(sh.alpha);
            double2 tau = cz(), scale = cz();
            double beta = ac.x;
            if (nrm[0] > 0.0 || ac.y != 0.0) {
                beta = -copysign(sqrt(fma(ac.x, ac.x, fma(ac.y, ac.y, nrm[0]))), ac.x);
                tau = make_double2((beta - ac.x) / beta, -ac.y / beta);
                const double dx = ac.x - beta, dy = ac.y;
                const double dd = 1.0 / fma(dx, dx, dy * dy);
                scale = make_double2(dx * dd, -dy * dd);
            }
#pragma unroll
            for (int k = 0; k < CPT; ++k) {
                const int c = b0 + tid + k * LQ_NT;
                if (c > j && c < m) x[T][k] = cmul(cconj(x[T][k]), scale);
            }
            // acc[v]: w_v = x[v, j:] u (v > T) and g_v = u_v^H u (v < T).
            double acc[2 * LQ_NB];
#pragma unroll
            for (int v = 0; v < 2 * LQ_NB; ++v) acc[v] = 0.0;
#pragma unroll
            for (int k = 0; k < CPT; ++k) {
                const int c = b0 + tid + k * LQ_NT;
                if (c < j || c >= m) continue;
                const double2 u = c == j ? one : x[T][k];
#pragma unroll
                for (int v = 0; v < LQ_NB; ++v) {
                    if (v == T) continue;
                    const double2 p = v > T ? cmul(x[v][k], u) : cmul(cconj(x[v][k]), u);
                    acc[2 * v] += p.x;
                    acc[2 * v + 1] += p.y;
                }
            }
            lq_block_sum32(acc, sh.redv, sh.finv, tid);
#pragma unroll
            for (int v = T + 1; v < LQ_NB; ++v) {
                const double2 coef = cmul(tau, make_double2(acc[2 * v], acc[2 * v + 1]));
#pragma unroll
                for (int k = 0; k < CPT; ++k) {
                    const int c = b0 + tid + k * LQ_NT;
                    if (c < j || c >= m) continue;
                    const double2 u = c == j ? one : x[T][k];
                    x[v][k] = csub(x[v][k], cmul(coef, cconj(u)));
                }
            }
            // Column T of the panel's T: row i (one thread each) is sum_{l >= i} T[i][l] z_l with
            // z_l = -tau g_l, summed in increasing l.
            if (tid < T) {
                const double2 mt = make_double2(-tau.x, -tau.y);
                double2 a = cz();
#pragma unroll
                for (int l = 0; l < T; ++l)
                    if (l >= tid) cmac(a, sh.T[tid][l], cmul(mt, make_double2(acc[2 * l], acc[2 * l + 1])));
                sh.T[tid][T] = a;
            }
            if (tid == 0) {
                sh.beta[T] = beta;
                sh.T[T][T] = tau;
            }
        }
        lq_step<CPT, T + 1>(x, sh, b0, j0, m, nbp, tid);
    }
}

template <int CPT>
__global__ __launch_bounds__(LQ_NT) void k_lq_panel(const LqJob* __restrict__ jobs, int b0, int j0) {
    const LqJob jb = jobs[blockIdx.x];
    if (j0 >= jb.n || jb.m - b0 > CPT * LQ_NT) return;
    const int nbp = min(LQ_NB, jb.n - j0);
    const int tid = threadIdx.x;
    __shared__ LqShared sh;
    double2 x[LQ_NB][CPT];
#pragma unroll
    for (int r = 0; r < LQ_NB; ++r)
#pragma unroll
        for (int k = 0; k < CPT; ++k) {
            const int c = b0 + tid + k * LQ_NT;
            x[r][k] = (r < nbp && c < jb.m) ? jb.X[(long)(j0 + r) * jb.ldx + c] : cz();
        }
    if (tid < LQ_NB * LQ_NB) sh.T[tid / LQ_NB][tid % LQ_NB] = cz();
    lq_step<CPT, 0>(x, sh, b0, j0, jb.m, nbp, tid);
    __syncthreads();
    const double2 one = make_double2(1.0, 0.0);
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
        const int c = b0 + tid + k * LQ_NT;
        if (c >= jb.m) continue;
        double2 u[LQ_NB];
#pragma unroll
        for (int t = 0; t < LQ_NB; ++t) {
            const int j = j0 + t;
            u[t] = (t >= nbp || c < j) ? cz() : (c == j ? one : x[t][k]);
            if (t < nbp) {
                jb.X[(long)j * jb.ldx + c] = c < j ? x[t][k] : (c == j ? make_double2(sh.beta[t], 0.0) : cz());
                jb.Ubc[(long)j * jb.m + c] = cconj(u[t]);
            }
        }
#pragma unroll
        for (int tp = 0; tp < LQ_NB; ++tp) {
            if (tp >= nbp) continue;
            double2 a = cz();
#pragma unroll
            for (int t = 0; t < LQ_NB; ++t)
                if (t <= tp) cmac(a, u[t], sh.T[t][tp]);
            jb.UTc[(long)(j0 + tp) * jb.m + c] = cconj(a);
        }
    }
    // The panel's T is the diagonal block of the block's T.
    if (tid < LQ_NB * LQ_NB) {
        const int i = tid / LQ_NB, k = tid % LQ_NB;
        if (i < nbp && k < nbp) jb.T64[(j0 - b0 + i) * LQ_SB + (j0 - b0 + k)] = sh.T[i][k];
    }
}

// Completes the block's T from its diagonal panel blocks and Gu = U^H U by merging compact-WY
// factors pairwise: for adjacent groups A, B of panels, T_AB = -T_A (U_A^H U_B) T_B (the
// upper-right block of T for Q_A Q_B); panels (0,1) and (2,3) first, then (01, 23). The
// strictly lower part is zero.
__global__ __launch_bounds__(256) void k_lq_tblock(const LqJob* __restrict__ jobs, int b0) {
    const LqJob jb = jobs[blockIdx.x];
    const int nbs = min(LQ_SB, jb.n - b0);
    if (nbs <= 0) return;
    __shared__ double2 T[LQ_SB][LQ_SB + 1];
    __shared__ double2 W[LQ_SB / 2][LQ_SB / 2 + 1];
    // Gu staged in LDS: the merges below read it along dependent accumulation chains.
    __shared__ double2 G[LQ_SB][LQ_SB + 1];
    const int tid = threadIdx.x;
    for (int e = tid; e < LQ_SB * LQ_SB; e += 256) {
        const int i = e / LQ_SB, k = e % LQ_SB;
        const bool diag_block = i / LQ_NB == k / LQ_NB && k >= i;
        T[i][k] = (i < nbs && k < nbs && diag_block) ? jb.T64[e] : cz();
        G[i][k] = (i < nbs && k < nbs) ? jb.Gu[e] : cz();
    }
    __syncthreads();
    for (int w = LQ_NB; w < LQ_SB; w *= 2) {
        // Groups [a, a + w) and [a + w, a + 2w) for a = 0, 2w, ...
        // W = Gu[A, B] T_B (w x w per pair), then T[A, B] = -T_A W.
        for (int a = 0; a + w < nbs; a += 2 * w) {
            const int bw = min(w, nbs - a - w);
            for (int e = tid; e < w * bw; e += 256) {
                const int i = e / bw, k = e % bw;
                double2 acc = cz();
                for (int l = 0; l <= k; ++l) cmac(acc, G[a + i][a + w + l], T[a + w + l][a + w + k]);
                W[i][k] = acc;
            }
            __syncthreads();
            for (int e = tid; e < w * bw; e += 256) {
                const int i = e / bw, k = e % bw;
                double2 acc = cz();
                for (int l = i; l < w; ++l) cmac(acc, T[a + i][a + l], W[l][k]);
                T[a + i][a + w + k] = make_double2(-acc.x, -acc.y);
            }
            __syncthreads();
        }
    }
    for (int e = tid; e < LQ_SB * LQ_SB; e += 256) jb.T64[e] = T[e / LQ_SB][e % LQ_SB];
}

// R = L^H (n x n, row stride n) from the lower triangle of the factored X.
__global__ __launch_bounds__(256) void k_lq_extract(const LqJob* __restrict__ jobs) {
    const LqJob jb = jobs[blockIdx.y];
    const long total = (long)jb.n * jb.n;
    for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
        const int i = (int)(e / jb.n), k = (int)(e - (long)i * jb.n);
        jb.R[e] = k >= i ? cconj(jb.X[(long)k * jb.ldx + i]) : cz();
    }
}

// X <- [J 0] (n x m, row stride ldx).
__global__ __launch_bounds__(256) void k_lq_init_m(const LqJob* __restrict__ jobs) {
    const LqJob jb = jobs[blockIdx.y];
    const long total = (long)jb.n * jb.m;
    for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
        const int i = (int)(e / jb.m), c = (int)(e - (long)i * jb.m);
        jb.X[(long)i * jb.ldx + c] = c < jb.n ? jb.J[(long)i * jb.n + c] : cz();
    }
}

// ------------------------------------------------------------------ Householder QR
// One workgroup per job; see HouseJob. Reflector convention v = x + (x0/|x0|) ||x|| e1,
// normalized (the reflection axis nalgebra computes), H = I - 2 v v^H.
__global__ __launch_bounds__(1024) void k_householder(const HouseJob* __restrict__ jobs) {
    const HouseJob jb = jobs[blockIdx.x];
    const int rows = jb.rows, r = jb.r;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    __shared__ double red[4];
    __shared__ double s_cn[kMaxHouseCols];
    __shared__ double s_tau[kMaxHouseCols];
    __shared__ double2 s_w[kMaxHouseCols];
    __shared__ double2 s_x0;
    __shared__ double s_alpha, s_vn;
    __shared__ int s_skip;
    __shared__ double2 s_part[16][64];  // per-wave partial column sums
    __shared__ double2 s_f[64];
    double2* __restrict__ Y = jb.Y;
    double2* __restrict__ V = jb.V;
    // Sixteen waves. The column norms give a column to one wave and the reflector norm runs on
    // the first four waves; the trailing update and T run lanes along rows (see there).
    // Original column norms.
    for (int k = wave; k < r; k += 16) {
        double acc = 0.0;
        for (int i = lane; i < rows; i += 64) {
            const double2 v = Y[(long)i * r + k];
            acc = fma(v.x, v.x, fma(v.y, v.y, acc));
        }
        acc = wave_sum(acc);
        if (lane == 0) s_cn[k] = sqrt(acc);
    }
    for (long e = tid; e < (long)rows * r; e += 1024) V[e] = cz();
    __syncthreads();
    for (int k = 0; k < r; ++k) {
        const int j = jb.pos[k];
        double acc = 0.0;
        if (tid < 256)
            for (int i = j + tid; i < rows; i += 256) {
                const double2 v = Y[(long)i * r + k];
                acc = fma(v.x, v.x, fma(v.y, v.y, acc));
            }
        acc = wave_sum(acc);
        if (lane == 0 && wave < 4) red[wave] = acc;
        __syncthreads();
        if (tid == 0) {
            const double a2 = ((red[0] + red[1]) + red[2]) + red[3];
            const double alpha = sqrt(a2);
            s_skip = !(alpha > 0.0) || alpha <= jb.tau * s_cn[k] || j >= rows;
            if (!s_skip) {
                const double2 x0 = Y[(long)j * r + k];
                const double x0n = hypot(x0.x, x0.y);
                const double2 ph = x0n > 0.0 ? make_double2(x0.x / x0n, x0.y / x0n) : make_double2(1.0, 0.0);
                const double2 nx0 = make_double2(x0.x + ph.x * alpha, x0.y + ph.y * alpha);
                // ||v||^2 = ||x||^2 - |x0|^2 + |x0 + ph alpha|^2
                const double vn2 = a2 - x0n * x0n + (nx0.x * nx0.x + nx0.y * nx0.y);
                s_x0 = nx0;
                s_vn = sqrt(vn2);
                s_alpha = alpha;
            }
            s_tau[k] = s_skip ? 0.0 : 2.0;
        }
        __syncthreads();
        if (s_skip) continue;
        const double inv = 1.0 / s_vn;
        for (int i = j + tid; i < rows; i += 1024) {
            const double2 x = (i == j) ? s_x0 : Y[(long)i * r + k];
            V[(long)i * r + k] = cscale(x, inv);
        }
        __syncthreads();
        // Trailing columns: Y[:, c] -= 2 v (v^H Y[:, c]). Lanes run along a row (coalesced:
        // Y and V are row-major with r columns), wave w takes rows j + w, j + w + 16, ...;
        // the sixteen wave partials of each column are added in wave order.
        for (int c0 = k + 1; c0 < r; c0 += 64) {
            const int c = c0 + lane;
            const bool act = c < r;
            double2 d = cz();
            if (act)
                for (int i = j + wave; i < rows; i += 16) ccmac(d, V[(long)i * r + k], Y[(long)i * r + c]);
            s_part[wave][lane] = d;
            __syncthreads();
            if (wave == 0) {
                double2 t = cz();
                for (int w = 0; w < 16; ++w) t = cadd(t, s_part[w][lane]);
                s_f[lane] = cscale(t, 2.0);
            }
            __syncthreads();
            if (act) {
                const double2 f = s_f[lane];
                for (int i = j + wave; i < rows; i += 16) {
                    const long at = (long)i * r + c;
                    Y[at] = csub(Y[at], cmul(V[(long)i * r + k], f));
                }
            }
            __syncthreads();
        }
    }
    // T (upper triangular): T[k][k] = tau_k, T[0:k, k] = -tau_k T[0:k, 0:k] (V[:,0:k]^H v_k).
    double2* T = jb.T;
    for (long e = tid; e < (long)r * r; e += 1024) T[e] = cz();
    __syncthreads();
    for (int k = 0; k < r; ++k) {
        if (s_tau[k] == 0.0) continue;
        // s_w[i] = V[:, i]^H v_k for i < k, laid out like the trailing update.
        for (int i0 = 0; i0 < k; i0 += 64) {
            const int i = i0 + lane;
            double2 d = cz();
            if (i < k)
                for (int row = wave; row < rows; row += 16) ccmac(d, V[(long)row * r + i], V[(long)row * r + k]);
            s_part[wave][lane] = d;
            __syncthreads();
            if (wave == 0 && i < k) {
                double2 t = cz();
                for (int w = 0; w < 16; ++w) t = cadd(t, s_part[w][lane]);
                s_w[i] = t;
            }
            __syncthreads();
        }
        if (tid < k) {
            double2 acc = cz();
            for (int l = tid; l < k; ++l) cmac(acc, T[(long)tid * r + l], s_w[l]);
            T[(long)tid * r + k] = cscale(acc, -s_tau[k]);
        }
        if (tid == 0) T[(long)k * r + k] = make_double2(s_tau[k], 0.0);
        __syncthreads();
    }
}

// ------------------------------------------------------------------ vector x site
// Left pass: workgroup (x, y, z) sums rows y*kVecMatChunk.. of v[a] site[a, s, o] of job z for
// 256 consecutive o (coalesced along the row); a second kernel adds the chunks in order. A
// job's output length is its site's r and its chunk count ceil(min(l, vlen) / kVecMatChunk).
// Right pass: one wave per output row o, coalesced along the row, wave reduction.
__global__ __launch_bounds__(256) void k_vecmat_left(const VecMatJob* __restrict__ jobs) {
    const VecMatJob jb = jobs[blockIdx.z];
    const int out_len = jb.r;
    const int o = blockIdx.x * 256 + threadIdx.x;
    if (o >= out_len) return;
    const int lim = jb.l < jb.vlen ? jb.l : jb.vlen;
    const int a0 = blockIdx.y * kVecMatChunk;
    if (blockIdx.y > 0 && a0 >= lim) return;
    const int a1 = a0 + kVecMatChunk < lim ? a0 + kVecMatChunk : lim;
    const long rs = 2L * jb.r;
    double2 acc = cz();
    for (int a = a0; a < a1; ++a) {
        const double2* s = jb.site + a * rs + o;
        double2 w = jb.phys == 1 ? s[jb.r] : s[0];
        if (jb.phys == 2) w = cadd(w, s[jb.r]);
        cmac(acc, jb.vin[a], w);
    }
    jb.part[(long)blockIdx.y * out_len + o] = acc;
}

__global__ __launch_bounds__(256) void k_vecmat_sum(const VecMatJob* __restrict__ jobs) {
    const VecMatJob jb = jobs[blockIdx.z];
    const int out_len = jb.r;
    const int o = blockIdx.x * 256 + threadIdx.x;
    if (o >= out_len) return;
    const int lim = jb.l < jb.vlen ? jb.l : jb.vlen;
    const int nchunk = (lim + kVecMatChunk - 1) / kVecMatChunk;
    double2 acc = cz();
    for (int c = 0; c < nchunk; ++c) acc = cadd(acc, jb.part[(long)c * out_len + o]);
    jb.vout[o] = acc;
}

__global__ __launch_bounds__(256) void k_vecmat_right(const VecMatJob* __restrict__ jobp, int out_len) {
    const VecMatJob jb = *jobp;
    const int o = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (o >= out_len) return;
    const int lim = jb.r < jb.vlen ? jb.r : jb.vlen;
    const double2* s = jb.site + (long)o * 2 * jb.r;
    double re = 0.0, im = 0.0;
    for (int b = lane; b < lim; b += 64) {
        double2 w = jb.phys == 1 ? s[jb.r + b] : s[b];
        if (jb.phys == 2) w = cadd(w, s[jb.r + b]);
        const double2 v = jb.vin[b];
        re = fma(w.x, v.x, fma(-w.y, v.y, re));
        im = fma(w.x, v.y, fma(w.y, v.x, im));
    }
    re = wave_sum(re);
    im = wave_sum(im);
    if (lane == 0) jb.vout[o] = make_double2(re, im);
}

// dst (cols x rows) = src (rows x cols)^H
__global__ __launch_bounds__(256) void k_conj_transpose(const double2* __restrict__ src,
                                                        double2* __restrict__ dst, int rows,
                                                        int cols) {
    __shared__ double2 tile[32][33];
    const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int i = ty; i < 32; i += 8) {
        const int r = r0 + i, c = c0 + tx;
        tile[i][tx] = (r < rows && c < cols) ? src[(long)r * cols + c] : cz();
    }
    __syncthreads();
    for (int i = ty; i < 32; i += 8) {
        const int c = c0 + i, r = r0 + tx;
        if (c < cols && r < rows) dst[(long)c * rows + r] = cconj(tile[tx][i]);
    }
}

__global__ void k_fill_int(int* p, int value, int count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = value;
}

__global__ void k_add_identity(double2* C, int rows, int cols, int ldc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < rows && i < cols) C[(long)i * ldc + i].x += 1.0;
}

// next[b*(2p) + ph*p + bp] = sum_a st[a*p + bp] * site[a, ph, b]
// ------------------------------------------------------------------ expectation values
// Site A (l, 2, r) and its Pauli-transformed copies B_t = sum_s' conj(P[s', t]) A_s' for
// P = X, Y, Z: X swaps the physical halves, Y gives (-i A_1, i A_0), Z gives (A_0, -A_1).
// With them one transfer step of <psi|P|psi> is E' = B_stack^H (E A_flat) (two GEMMs).
__global__ __launch_bounds__(256) void k_pauli_site(const double2* __restrict__ A, double2* __restrict__ bx,
                                                    double2* __restrict__ by, double2* __restrict__ bz,
                                                    int l, int r) {
    const long half = (long)l * r;
    for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < half; e += (long)gridDim.x * 256) {
        const long a = e / r, b = e % r;
        const long i0 = a * 2 * r + b, i1 = i0 + r;
        const double2 v0 = A[i0], v1 = A[i1];
        bx[i0] = v1;
        bx[i1] = v0;
        by[i0] = make_double2(v1.y, -v1.x);   // -i * v1
        by[i1] = make_double2(-v0.y, v0.x);   //  i * v0
        bz[i0] = v0;
        bz[i1] = make_double2(-v1.x, -v1.y);
    }
}

// out = sum_i E[i] conj(R[i]) per job, one workgroup per job, fixed-order reduction.
__global__ __launch_bounds__(256) void k_env_close(const CloseJob* __restrict__ jobs) {
    const CloseJob jb = jobs[blockIdx.x];
    __shared__ double2 red[256];
    double2 acc = cz();
    for (long i = threadIdx.x; i < jb.count; i += 256) ccmac(acc, jb.R[i], jb.E[i]);
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] = cadd(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) *jb.out = red[0];
}

// ------------------------------------------------------------------ MPO x MPS, operator sites
// One workgroup per (target, s): the bilinear form Lv . site[:, s, :] . Rv, summed per thread
// over a strided share of (a, b) and then over the threads in a fixed tree.
__global__ __launch_bounds__(256) void k_site_amps(const SiteAmpJob* __restrict__ jobs) {
    const SiteAmpJob jb = jobs[blockIdx.x];
    const int s = blockIdx.y, tid = threadIdx.x;
    const int la = min(jb.l, jb.llen);
    const long total = (long)la * jb.r;
    double2 acc = cz();
    for (long e = tid; e < total; e += 256) {
        const int a = (int)(e / jb.r), b = (int)(e - (long)a * jb.r);
        const double2 w = b < jb.rlen ? jb.Rv[b] : make_double2(1.0, 0.0);
        cmac(acc, cmul(jb.Lv[a], jb.site[((long)a * 2 + s) * jb.r + b]), w);
    }
    __shared__ double2 red[256];
    red[tid] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (tid < w) red[tid] = cadd(red[tid], red[tid + w]);
        __syncthreads();
    }
    if (tid == 0) jb.out[s] = red[0];
}

__global__ __launch_bounds__(256) void k_overlap_site(const double2* __restrict__ E, const double2* __restrict__ K,
                                                      const double2* __restrict__ B, double2* __restrict__ En,
                                                      int kl, int kr, int bl, int br) {
    __shared__ double2 Es[OVERLAP_ENV_MAX];
    const int tid = threadIdx.x;
    if (tid < kl * bl) Es[tid] = E[tid];
    __syncthreads();
    if (tid >= kr * br) return;
    const int ap = tid / br, bp = tid - ap * br;
    double2 acc = cz();
    for (int a = 0; a < kl; ++a)
        for (int b = 0; b < bl; ++b) {
            const double2 e = Es[a * bl + b];
            for (int s = 0; s < 2; ++s) cmacc(acc, cmul(e, K[(a * 2 + s) * kr + ap]), B[(b * 2 + s) * br + bp]);
        }
    En[tid] = acc;
}

__global__ __launch_bounds__(256) void k_mpo_site(const double2* __restrict__ A, const double2* __restrict__ W,
                                                  double2* __restrict__ out, int l, int r, int Dl, int Dr,
                                                  int w0, int wn, int v0, int vn) {
    const long total = (long)l * wn * 2 * r * vn;
    const long ro = (long)r * vn;
    for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
        const long row = e / (2 * ro), rem = e % (2 * ro);
        const int t = (int)(rem / ro);
        const long col = rem % ro;
        const int a = (int)(row / wn), w = w0 + (int)(row % wn);
        const int b = (int)(col / vn), v = v0 + (int)(col % vn);
        const double2* wv = W + (((long)w * Dr + v) * 2 + t) * 2;
        const double2 x0 = A[(long)a * 2 * r + b], x1 = A[(long)a * 2 * r + r + b];
        double2 acc = cz();
        cmac(acc, wv[0], x0);
        cmac(acc, wv[1], x1);
        out[e] = acc;
    }
}

// ------------------------------------------------------------------ adjoint sweep
// 2x2 correlation of a site cotangent with a site: M[s][t] = sum_{a,b} conj(C[a,s,b]) S[a,t,b],
// one workgroup per job, per-thread partial sums over a strided share of (a, b), then a fixed
// tree over the threads.
__global__ __launch_bounds__(256) void k_site_corr(const SiteCorrJob* __restrict__ jobs) {
    const SiteCorrJob jb = jobs[blockIdx.x];
    const int tid = threadIdx.x;
    const long total = (long)jb.l * jb.r;
    double2 m00 = cz(), m01 = cz(), m10 = cz(), m11 = cz();
    for (long e = tid; e < total; e += 256) {
        const long a = e / jb.r, b = e - a * jb.r;
        const long o0 = (a * 2) * jb.r + b, o1 = o0 + jb.r;
        const double2 c0 = jb.C[o0], c1 = jb.C[o1], s0 = jb.S[o0], s1 = jb.S[o1];
        ccmac(m00, c0, s0);
        ccmac(m01, c0, s1);
        ccmac(m10, c1, s0);
        ccmac(m11, c1, s1);
    }
    __shared__ double2 red[4][256];
    red[0][tid] = m00;
    red[1][tid] = m01;
    red[2][tid] = m10;
    red[3][tid] = m11;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (tid < w)
#pragma unroll
            for (int k = 0; k < 4; ++k) red[k][tid] = cadd(red[k][tid], red[k][tid + w]);
        __syncthreads();
    }
    if (tid < 4) jb.out[tid] = red[tid][0];
}

// Theta' -> theta cotangent through the two-qubit gate, in place: T[a,s,t,b] <-
// sum_{s',t'} conj(U4[(s',t')][(s,t)]) T[a,s',t',b] on a (2dl x 2dr) row-major matrix.
__global__ __launch_bounds__(256) void k_apply4_adj(double2* __restrict__ T, int dl, int dr, U4Arg u) {
    const long total = (long)dl * dr;
    for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
        const long a = e / dr, b = e - a * dr;
        const long w = 2L * dr;
        double2 v[4];
#pragma unroll
        for (int st = 0; st < 4; ++st) v[st] = T[(a * 2 + (st >> 1)) * w + (st & 1) * dr + b];
#pragma unroll
        for (int st = 0; st < 4; ++st) {
            double2 acc = cz();
#pragma unroll
            for (int sp = 0; sp < 4; ++sp) ccmac(acc, u.m[sp * 4 + st], v[sp]);
            T[(a * 2 + (st >> 1)) * w + (st & 1) * dr + b] = acc;
        }
    }
}

// Coefficients of the truncated-split adjoint (see SvdAdjJob). Entry (x, y) of C:
//   x == y, x kept as i:  P/s - conj(XV[x,i]) / s^3,  P = sbar + (i/2)(Im JU[x,i] - Im XV[x,i]/s)/s,
//                         sbar = Re(JU[x,i] + XV[x,i]/s) / (2s),  s = sigma_x;
//   x != y, both kept:    conj(al) / sigma_y - conj(XV[y,ix]) / (sigma_x sigma_y^2) with
//                         al = (a - b) / (2 (sigma_y + sigma_x)) + (a + b) / (2 (sigma_y - sigma_x)),
//                         a = conj(JU[x,iy]) - JU[y,ix], b = conj(JV[x,iy]) - JV[y,ix],
//                         JV[j,i] = XV[j,i] / sigma_j (the second term at its limit for equal values);
//   otherwise: [y kept as i]  (JU[x,i] s + XV[x,i]) / ((s^2 - sigma_x^2) s),  s = sigma_y
//         + [x kept as i]  conj(JU[y,i]) / (s^2 - sigma_y^2)
//                          + [sigma_y^2 > floor2] conj(XV[y,i]) / (s (s^2 - sigma_y^2)),  s = sigma_x,
// with JU[j,i] = JUraw[j,i] sqrt(s_i) and XV[j,i] = XVraw[j,i] sqrt(s_i) for kept row i; and
// E[x,i] = 1/sqrt(sigma_x) where row x is kept as i.
__global__ __launch_bounds__(256) void k_svd_adj_coef(SvdAdjJob jb) {
    const int x = blockIdx.y * 16 + (threadIdx.x >> 4), y = blockIdx.x * 16 + (threadIdx.x & 15);
    if (x < jb.p && y < jb.c) {
        const int i = jb.inv[x];
        jb.E[(long)x * jb.c + y] = make_double2(i == y ? 1.0 / sqrt(jb.sig[x]) : 0.0, 0.0);
    }
    if (x >= jb.p || y >= jb.p) return;
    double2 val = cz();
    if (x == y) {
        const int i = jb.inv[x];
        if (i >= 0) {
            const double s = jb.sig[x], rs = sqrt(s);
            const double2 ju = cscale(jb.JU[(long)x * jb.c + i], rs), xv = cscale(jb.XV[(long)x * jb.c + i], rs);
            const double sbar = (ju.x + xv.x / s) / (2.0 * s);
            const double pim = 0.5 * (ju.y - xv.y / s) / s;
            const double s3 = s * s * s;
            val = make_double2(sbar / s - xv.x / s3, pim / s + xv.y / s3);
        }
    } else if (jb.inv[x] >= 0 && jb.inv[y] >= 0) {
        // Both rows kept: the pair's coefficient split into its 1/(s_y + s_x) part and its
        // 1/(s_y - s_x) part, which for (near-)equal values is replaced by its limit (the
        // cotangents of a rotation inside a degenerate pair cancel: a + b -> 0).
        const int ix = jb.inv[x], iy = jb.inv[y];
        const double sx = jb.sig[x], sy = jb.sig[y];
        const double2 jux = cscale(jb.JU[(long)x * jb.c + iy], sqrt(sy)), juy = cscale(jb.JU[(long)y * jb.c + ix], sqrt(sx));
        const double2 xvx = cscale(jb.XV[(long)x * jb.c + iy], sqrt(sy)), xvy = cscale(jb.XV[(long)y * jb.c + ix], sqrt(sx));
        const double2 jvx = cscale(xvx, 1.0 / sx), jvy = cscale(xvy, 1.0 / sy);
        const double2 a = csub(cconj(jux), juy), b = csub(cconj(jvx), jvy);
        const double dm = sy - sx, dp = sy + sx;
        double2 al = cscale(csub(a, b), 0.5 / dp);
        if (fabs(dm) > 1e-8 * fmax(sx, sy))
            al = cadd(al, cscale(cadd(a, b), 0.5 / dm));
        else
            al = cadd(al, cscale(cadd(cadd(cconj(jux), jvy), cadd(juy, cconj(jvx))), 0.5 / dp));
        val = csub(cscale(cconj(al), 1.0 / sy), cscale(cconj(xvy), 1.0 / (sx * sy * sy)));
    } else {
        const int iy = jb.inv[y], ix = jb.inv[x];
        if (iy >= 0) {
            const double s = jb.sig[y], sx = jb.sig[x], rs = sqrt(s);
            const double d = (s - sx) * (s + sx);
            if (d != 0.0) {
                const double2 ju = cscale(jb.JU[(long)x * jb.c + iy], rs), xv = cscale(jb.XV[(long)x * jb.c + iy], rs);
                val = cadd(val, cscale(cadd(cscale(ju, s), xv), 1.0 / (d * s)));
            }
        }
        if (ix >= 0) {
            const double s = jb.sig[x], sy = jb.sig[y], rs = sqrt(s);
            const double d = (s - sy) * (s + sy);
            if (d != 0.0) {
                const double2 ju = cscale(jb.JU[(long)y * jb.c + ix], rs);
                val = cadd(val, cscale(cconj(ju), 1.0 / d));
                if (sy * sy > jb.floor2) {
                    const double2 xv = cscale(jb.XV[(long)y * jb.c + ix], rs);
                    val = cadd(val, cscale(cconj(xv), 1.0 / (s * d)));
                }
            }
        }
    }
    jb.C[(long)x * jb.p + y] = val;
}

// ------------------------------------------------------------------ MPS sampler step
// One wave per shot: p_s = Re sum_b conj(W_s[b]) Z_s[b] = <v A_s| R |v A_s>, the unnormalized
// probability of outcome s given the shot's earlier outcomes; s = 0 when u < p_0 / (p_0 + p_1).
__global__ __launch_bounds__(256) void k_sample_step(const double2* __restrict__ W,
                                                     const double2* __restrict__ Z, int r,
                                                     long shots, const double* __restrict__ u,
                                                     long ustride, int bit,
                                                     unsigned long long* __restrict__ bits,
                                                     double2* __restrict__ vnext) {
    const long k = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (k >= shots) return;
    const double2* w = W + k * 2 * r;
    const double2* z0 = Z + k * r;
    const double2* z1 = Z + (shots + k) * r;
    double p0 = 0.0, p1 = 0.0;
    for (int b = lane; b < r; b += 64) {
        p0 = fma(w[b].x, z0[b].x, fma(w[b].y, z0[b].y, p0));
        p1 = fma(w[r + b].x, z1[b].x, fma(w[r + b].y, z1[b].y, p1));
    }
    p0 = fmax(wave_sum(p0), 0.0);
    p1 = fmax(wave_sum(p1), 0.0);
    const double tot = p0 + p1;
    const int s = (tot > 0.0 && u[k * ustride] < p0 / tot) ? 0 : 1;
    const double ps = s ? p1 : p0;
    const double scale = ps > 0.0 ? 1.0 / sqrt(ps) : 0.0;
    for (int b = lane; b < r; b += 64) {
        const double2 v = w[s * r + b];
        vnext[k * r + b] = make_double2(v.x * scale, v.y * scale);
    }
    if (lane == 0 && s) bits[k] |= 1ull << bit;
}

__global__ __launch_bounds__(256) void k_statevector_step(const double2* __restrict__ st,
                                                          const double2* __restrict__ site,
                                                          double2* __restrict__ next, int l,
                                                          int r, long prefix) {
    const long total = (long)r * 2 * prefix;
    for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
        const long b = e / (2 * prefix), rem = e - b * 2 * prefix;
        const long ph = rem / prefix, bp = rem - ph * prefix;
        double2 acc = cz();
        for (int a = 0; a < l; ++a) cmac(acc, st[a * prefix + bp], site[(long)a * 2 * r + ph * r + b]);
        next[e] = acc;
    }
}

inline int blocks_for(long n, int per = 256, int cap = 4096) {
    long b = (n + per - 1) / per;
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return (int)b;
}

}  // namespace

void launch_one_qubit(const OneQubitJob* jobs, int njobs, int max_elems, hipStream_t s) {
    dim3 grid(blocks_for(max_elems, 256, 1024), njobs);
    hipLaunchKernelGGL(k_one_qubit, grid, dim3(256), 0, s, jobs);
}

void launch_merge(const MergeJob* jobs, int njobs, int max_dl, int max_dr, hipStream_t s) {
    dim3 grid((max_dr + MA - 1) / MA, (max_dl + MA - 1) / MA, njobs);
    hipLaunchKernelGGL(k_merge, grid, dim3(256), 0, s, jobs);
}

void launch_sumsq(const double2* const* ptrs, const long* counts, double* partial, int njobs,
                  int nblk, hipStream_t s) {
    hipLaunchKernelGGL(k_sumsq, dim3(nblk, njobs), dim3(256), 0, s, ptrs, counts, partial, nblk);
}

void launch_sumsq_final(const double* partial, double* out, int njobs, int nblk, hipStream_t s) {
    hipLaunchKernelGGL(k_sumsq_final, dim3((njobs + 63) / 64), dim3(64), 0, s, partial, out,
                       njobs, nblk);
}

void launch_dust(double2* const* ptrs, const long* counts, const double* fro2, double rel,
                 int njobs, long max_count, hipStream_t s) {
    hipLaunchKernelGGL(k_dust, dim3(blocks_for(max_count, 256, 2048), njobs), dim3(256), 0, s,
                       ptrs, counts, fro2, rel);
}

void launch_identity(double2* const* ptrs, const int* dims, int njobs, int max_dim,
                     hipStream_t s) {
    hipLaunchKernelGGL(k_identity, dim3(blocks_for((long)max_dim * max_dim, 256, 2048), njobs),
                       dim3(256), 0, s, ptrs, dims);
}

void launch_jacobi_round(const JacobiJob* jobs, int njobs, int max_pairs, int round, double tol,
                         double nu2, int inner_cap, int* rotated, const int* done, unsigned long long* stats,
                         unsigned long long* offmax, bool wide, hipStream_t s) {
    if (wide)
        hipLaunchKernelGGL(k_jacobi_round<512>, dim3(max_pairs, njobs), dim3(512), 0, s, jobs, round, tol,
                           nu2, inner_cap, rotated, done, stats, offmax);
    else
        hipLaunchKernelGGL(k_jacobi_round<256>, dim3(max_pairs, njobs), dim3(256), 0, s, jobs, round, tol,
                           nu2, inner_cap, rotated, done, stats, offmax);
}

void launch_jacobi_round_lo(const JacobiLoJob* jobs, int njobs, int max_pairs, int round, float tol,
                            double nu2, float rel_floor, int inner_cap, int* rotated, const int* done, unsigned long long* stats,
                            unsigned long long* offmax, bool wide, bool bf16_gram, hipStream_t s) {
#define WQC_LO(NT, GB)                                                                                        \
    hipLaunchKernelGGL((k_jacobi_round_lo<NT, GB>), dim3(max_pairs, njobs), dim3(NT), 0, s, jobs, round, tol, \
                       nu2, rel_floor, inner_cap, rotated, done, stats, offmax)
    if (wide) {
        if (bf16_gram)
            WQC_LO(512, true);
        else
            WQC_LO(512, false);
    } else {
        if (bf16_gram)
            WQC_LO(256, true);
        else
            WQC_LO(256, false);
    }
#undef WQC_LO
}

void launch_to_f32(const double2* const* src, float2* const* dst, const long* counts, int njobs,
                   long max_count, hipStream_t s) {
    hipLaunchKernelGGL(k_to_f32, dim3(blocks_for(max_count, 256, 2048), njobs), dim3(256), 0, s, src,
                       dst, counts);
}

void launch_to_f64(const float2* const* src, double2* const* dst, const long* counts, int njobs,
                   long max_count, hipStream_t s) {
    hipLaunchKernelGGL(k_to_f64, dim3(blocks_for(max_count, 256, 2048), njobs), dim3(256), 0, s, src,
                       dst, counts);
}

void launch_copy_batch(const CopyJob* jobs, int njobs, long max_count, hipStream_t s) {
    if (njobs <= 0) return;
    hipLaunchKernelGGL(k_copy_batch, dim3(blocks_for(max_count, 256, 2048), njobs), dim3(256), 0, s, jobs);
}

void launch_probe_residual(const ProbeJob* jobs, int njobs, hipStream_t s) {
    hipLaunchKernelGGL(k_probe_residual, dim3(njobs), dim3(256), 0, s, jobs);
}

void launch_row_norms(const JacobiJob* jobs, int njobs, int max_n, double* sigma,
                      const long* sigma_off, hipStream_t s) {
    hipLaunchKernelGGL(k_row_norms, dim3((max_n + 3) / 4, njobs), dim3(256), 0, s, jobs, sigma,
                       sigma_off);
}

void launch_pauli_site(const double2* A, double2* bx, double2* by, double2* bz, int l, int r,
                       hipStream_t s) {
    const long half = (long)l * r;
    hipLaunchKernelGGL(k_pauli_site, dim3(blocks_for(half, 256, 4096)), dim3(256), 0, s, A, bx, by, bz, l, r);
}

void launch_mpo_site(const double2* A, const double2* W, double2* out, int l, int r, int Dl, int Dr,
                     int w0, int wn, int v0, int vn, hipStream_t s) {
    const long total = (long)l * wn * 2 * r * vn;
    hipLaunchKernelGGL(k_mpo_site, dim3(blocks_for(total, 256, 4096)), dim3(256), 0, s, A, W, out, l, r, Dl,
                       Dr, w0, wn, v0, vn);
}

void launch_sample_step(const double2* W, const double2* Z, int r, long shots, const double* u,
                        long ustride, int bit, unsigned long long* bits, double2* vnext,
                        hipStream_t s) {
    hipLaunchKernelGGL(k_sample_step, dim3((unsigned)((shots + 3) / 4)), dim3(256), 0, s, W, Z, r,
                       shots, u, ustride, bit, bits, vnext);
}

void launch_site_amps(const SiteAmpJob* jobs, int njobs, hipStream_t s) {
    hipLaunchKernelGGL(k_site_amps, dim3(njobs, 2), dim3(256), 0, s, jobs);
}

void launch_overlap_site(const double2* E, const double2* K, const double2* B, double2* En, int kl, int kr,
                         int bl, int br, hipStream_t s) {
    hipLaunchKernelGGL(k_overlap_site, dim3(1), dim3(256), 0, s, E, K, B, En, kl, kr, bl, br);
}

void launch_env_close(const CloseJob* jobs, int njobs, hipStream_t s) {
    hipLaunchKernelGGL(k_env_close, dim3(njobs), dim3(256), 0, s, jobs);
}

void launch_gemm(const GemmJob* jobs, int njobs, int max_m, int max_n, hipStream_t s, int tile) {
    if (tile == 32) {
        dim3 grid((max_n + 31) / 32, (max_m + 31) / 32, njobs);
        hipLaunchKernelGGL((k_gemm_mfma<false, 32>), grid, dim3(256), 0, s, jobs, 1);
        return;
    }
    dim3 grid((max_n + MG - 1) / MG, (max_m + MG - 1) / MG, njobs);
    hipLaunchKernelGGL((k_gemm_mfma<false, MG>), grid, dim3(256), 0, s, jobs, 1);
}

void launch_gemm_splitk(const GemmJob* jobs, int njobs, int max_m, int max_n, int splits, hipStream_t s, int tile) {
    if (tile == 32) {
        dim3 grid((max_n + 31) / 32, (max_m + 31) / 32, njobs * splits);
        hipLaunchKernelGGL((k_gemm_mfma<true, 32>), grid, dim3(256), 0, s, jobs, splits);
    } else {
        dim3 grid((max_n + MG - 1) / MG, (max_m + MG - 1) / MG, njobs * splits);
        hipLaunchKernelGGL((k_gemm_mfma<true, MG>), grid, dim3(256), 0, s, jobs, splits);
    }
    hipLaunchKernelGGL(k_gemm_splitk_reduce, dim3(blocks_for((long)max_m * max_n, 256, 256), njobs), dim3(256), 0,
                       s, jobs, splits);
}

void launch_factor(const FactorJob* jobs, int njobs, long max_left, long max_right,
                   hipStream_t s) {
    long mx = max_left > max_right ? max_left : max_right;
    hipLaunchKernelGGL(k_factor, dim3(blocks_for(mx, 256, 2048), njobs, 2), dim3(256), 0, s, jobs);
}

void launch_householder(const HouseJob* jobs, int njobs, hipStream_t s) {
    hipLaunchKernelGGL(k_householder, dim3(njobs), dim3(1024), 0, s, jobs);
}

void launch_lq_panel(const LqJob* jobs, int njobs, int b0, int j0, int max_cols, hipStream_t s) {
    if (max_cols <= LQ_NT)
        hipLaunchKernelGGL(k_lq_panel<1>, dim3(njobs), dim3(LQ_NT), 0, s, jobs, b0, j0);
    else
        hipLaunchKernelGGL(k_lq_panel<2>, dim3(njobs), dim3(LQ_NT), 0, s, jobs, b0, j0);
}

void launch_lq_tblock(const LqJob* jobs, int njobs, int b0, hipStream_t s) {
    hipLaunchKernelGGL(k_lq_tblock, dim3(njobs), dim3(256), 0, s, jobs, b0);
}

void launch_lq_extract(const LqJob* jobs, int njobs, int max_n, hipStream_t s) {
    hipLaunchKernelGGL(k_lq_extract, dim3(blocks_for((long)max_n * max_n, 256, 1024), njobs), dim3(256), 0, s,
                       jobs);
}

void launch_lq_init_m(const LqJob* jobs, int njobs, long max_elems, hipStream_t s) {
    hipLaunchKernelGGL(k_lq_init_m, dim3(blocks_for(max_elems, 256, 1024), njobs), dim3(256), 0, s, jobs);
}

void launch_vecmat(const VecMatJob* job, int out_len, int in_len, int right, hipStream_t s) {
    if (right) {
        hipLaunchKernelGGL(k_vecmat_right, dim3((out_len + 3) / 4), dim3(256), 0, s, job, out_len);
        return;
    }
    launch_vecmat_batch(job, 1, out_len, in_len, s);
}

void launch_vecmat_batch(const VecMatJob* jobs, int njobs, int max_out, int max_in, hipStream_t s) {
    if (njobs <= 0) return;
    const int nchunk = max_in > 0 ? (max_in + kVecMatChunk - 1) / kVecMatChunk : 1;
    hipLaunchKernelGGL(k_vecmat_left, dim3((max_out + 255) / 256, nchunk, njobs), dim3(256), 0, s, jobs);
    hipLaunchKernelGGL(k_vecmat_sum, dim3((max_out + 255) / 256, 1, njobs), dim3(256), 0, s, jobs);
}

void launch_conj_transpose(const double2* src, double2* dst, int rows, int cols, hipStream_t s) {
    dim3 grid((cols + 31) / 32, (rows + 31) / 32);
    hipLaunchKernelGGL(k_conj_transpose, grid, dim3(256), 0, s, src, dst, rows, cols);
}

void launch_fill_int(int* p, int value, int count, hipStream_t s) {
    if (count <= 0) return;
    hipLaunchKernelGGL(k_fill_int, dim3((count + 255) / 256), dim3(256), 0, s, p, value, count);
}

void launch_add_identity(double2* C, int rows, int cols, int ldc, hipStream_t s) {
    int n = rows < cols ? rows : cols;
    if (n <= 0) return;
    hipLaunchKernelGGL(k_add_identity, dim3((n + 255) / 256), dim3(256), 0, s, C, rows, cols, ldc);
}

void launch_statevector_step(const double2* st, const double2* site, double2* next, int l,
                             int r, long prefix, hipStream_t s) {
    long total = (long)r * 2 * prefix;
    hipLaunchKernelGGL(k_statevector_step, dim3(blocks_for(total, 256, 8192)), dim3(256), 0, s,
                       st, site, next, l, r, prefix);
}

void launch_site_corr(const SiteCorrJob* jobs, int njobs, hipStream_t s) {
    if (njobs <= 0) return;
    hipLaunchKernelGGL(k_site_corr, dim3(njobs), dim3(256), 0, s, jobs);
}

void launch_apply4_adj(double2* T, int dl, int dr, const double2 u4[16], hipStream_t s) {
    U4Arg u;
    for (int k = 0; k < 16; ++k) u.m[k] = u4[k];
    hipLaunchKernelGGL(k_apply4_adj, dim3(blocks_for((long)dl * dr, 256, 4096)), dim3(256), 0, s, T, dl, dr, u);
}

void launch_svd_adj_coef(const SvdAdjJob& job, hipStream_t s) {
    const int cols = job.p > job.c ? job.p : job.c;
    dim3 grid((cols + 15) / 16, (job.p + 15) / 16);
    hipLaunchKernelGGL(k_svd_adj_coef, grid, dim3(256), 0, s, job);
}

}  // namespace wqc

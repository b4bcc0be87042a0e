// kernels.h — device job descriptors and launch wrappers of the gfx950 MPS kernels.
//
// Complex numbers are double2 (x = re, y = im). Matrices are row-major. Every launch
// wrapper takes a device-resident job table (one entry per independent job of a
// layer) and enqueues on the given stream; none of them synchronizes.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace wqc {

// One-qubit update of one site (l, 2, r) in place: site[a,s',b] = sum_s U[s'][s] site[a,s,b].
// One device-to-device copy of n doubles.
struct CopyJob {
    const double* src;
    double* dst;
    long n;
    long pad;
};

struct OneQubitJob {
    double2* site;
    int l, r;
    double2 u[4];  // U[0][0], U[0][1], U[1][0], U[1][1]
};

// theta' = U4 . (L . R). L is (2dl x bond), R is (bond x 2dr), both row-major views of
// sites (dl,2,bond) and (bond,2,dr). U4[(s',t')][(s,t)] is the two-qubit unitary with the
// control ordering already resolved. out is (2dl x 2dr) row-major, or its conjugate
// transpose (2dr x 2dl) when transpose != 0.
struct MergeJob {
    const double2* L;
    const double2* R;
    double2* out;
    int dl, bond, dr, transpose;
    double2 u4[16];
};

// Jacobi working matrix X (n x m, row stride ldx, n <= m): rows are rotated in place
// until mutually orthogonal; J (n x n, row stride n) accumulates the rotations (X_final =
// J X_orig) when non-null.
struct JacobiJob {
    double2* X;
    double2* J;
    int n, m, ldx, pad;
};

// Mixed-precision Jacobi job: rows X in complex64 (n x m, row stride ldx), rotation J
// (n x n, complex128) accumulated so that J stays unitary to FP64 precision.
struct JacobiLoJob {
    float2* X;
    float2* J;  // n x n, complex64
    int n, m, ldx, pad;
};

// Batched complex GEMM: C = alpha * op(A) * op(B) + beta * C, op in {N, C} (C = conjugate
// transpose). M x N result, inner dimension K. Leading dimensions are row strides of the
// stored matrices.
// With split-K (launch_gemm_splitk) every job's K range is cut into `splits` slices, slice s
// writes its raw product into part + s * M * N (row-major M x N), and a second kernel sums
// the slices in order and applies alpha and beta: more workgroups for the skinny, long-K
// products of the LQ, with a fixed summation order.
struct GemmJob {
    const double2* A;
    const double2* B;
    double2* C;
    int M, N, K;
    int lda, ldb, ldc;
    int opA, opB;  // 0 = N, 1 = C
    double2 alpha, beta;
    double2* part;  // split-K partials (splits * M * N), or null
};

// Factor writes of the split.
//   mode 0 (exact, X = theta', n = rows): left[r,k] = conj(J[p_k, r]) sqrt(s),
//                                          right[k,c] = X[p_k, c] / sqrt(s)
//   mode 1 (exact, X = theta'^H, n = cols): left[r,k] = conj(X[p_k, r]) / sqrt(s),
//                                           right[k,c] = J[p_k, c] sqrt(s)
//   mode 2 (range finder): right[k,c] = X[p_k, c] / sqrt(s);
//                          left (as W, ell x chi): W[t,k] = conj(J[p_k, t]) sqrt(s)
//   mode 3 (range finder, LQ-preconditioned): left (as W, ldx x chi) as in mode 1,
//                          right as in mode 1
// p_k = perm[k], s = sigma[p_k]; s == 0 writes zeros. After an LQ-preconditioned Jacobi
// the roles swap: X holds the rotated rows of R (norms s) and J the rows of [J 0] Q^H.
struct FactorJob {
    const double2* X;
    const double2* J;
    double2* left;
    double2* right;
    const int* perm;
    const double* sigma;
    int rows, cols, chi, ldx, ldj, mode, pad0, pad1;
};

// Householder LQ of a Jacobi working matrix X (n x m, row stride ldx, n <= m):
// X Q_1 Q_2 ... Q_p = [L 0], L lower triangular (n x n), each Q_i = I - U_i T_i U_i^H the
// compact WY form of a block of LQ_SB rows' reflectors (LAPACK zlarfg / zlarft forward-
// columnwise conventions). A block is factored as LQ_SB / LQ_NB register-resident panels:
// k_lq_panel leaves L's rows in X, writes the panel's rows of Ubc = U^H (unit diagonal,
// zeros to its left, from the block's first column on) and of UTc = (U T_panel)^H, and its
// T_panel into the diagonal of the block's T (T64, LQ_SB x LQ_SB); the block's later rows take
// X_rows -= (X_rows UTc^H) Ubc (two GEMMs). Then Gu = Ubc Ubc^H (GEMM), k_lq_tblock completes
// T64 (T[0:j, j] = -tau_j T[0:j, 0:j] Gu[0:j, j]), and the block's UTc = T64^H Ubc and
// UThc = T64 Ubc (GEMMs) serve the rows below, X_rest -= (X_rest UTc^H) Ubc. After the last
// block R = L^H (launch_lq_extract) is the Jacobi matrix; with J R = Y (rows orthogonal),
// X = Y^H M where M = [J 0] Q_p^H ... Q_1^H has orthonormal rows (launch_lq_init_m writes
// [J 0] into X, then per block in reverse M <- M - (M UThc^H) Ubc).
struct LqJob {
    double2* X;
    double2* Ubc;      // n x m
    double2* UTc;      // n x m
    double2* UThc;     // n x m
    double2* R;        // n x n
    const double2* J;  // n x n
    double2* T64;      // LQ_SB x LQ_SB (the current block's T)
    const double2* Gu; // LQ_SB x LQ_SB (the current block's U^H U)
    int n, m, ldx, pad;
};
constexpr int LQ_NB = 16;
constexpr int LQ_SB = 64;
constexpr int LQ_NT = 512;
constexpr int LQ_MAX_COLS = 2 * LQ_NT;

// Rank-thresholded Householder QR of the independent columns of Y (range finder).
// Y (rows x r, row stride r) is overwritten. Column k sits at position pos[k] of the
// full ell-column matrix; its reflector acts on rows >= pos[k]. V (rows x r) receives
// unit reflector vectors (zero for skipped columns), T (r x r) the compact-WY factor of
// H_pos[0] ... H_pos[r-1] = I - V T V^H.
constexpr int kMaxHouseCols = 1024;
struct HouseJob {
    double2* Y;
    double2* V;
    double2* T;
    const int* pos;
    int rows, r, pad0, pad1;
    double tau;
};

// v_next[b] = sum_a v[a] * w(a, b) over one site (l, 2, r), for a < min(l, vlen), where
// w = site[a,0,b] (phys = 0), site[a,1,b] (phys = 1) or their sum (phys = 2). With
// right != 0 the contraction runs from the right: v_next[a] = sum_b w(a,b) v[b].
// v' = v . site[:, s, :] (right = 0, length r) or site[:, s, :] . v (right = 1, length l), with
// s = phys or the sum over both (phys = 2); `part` is scratch for the left pass's row chunks.
struct VecMatJob {
    const double2* site;
    const double2* vin;
    double2* vout;
    int l, r, vlen, phys, right, pad;
    double2* part;
};

// Site amplitudes of one target from its left and right environments:
// out[s] = sum_{a < min(l, llen), b < r} Lv[a] site[a, s, b] (b < rlen ? Rv[b] : 1), s = 0, 1.
struct SiteAmpJob {
    const double2* Lv;
    const double2* site;
    const double2* Rv;
    double2* out;  // 2 values
    int llen, l, r, rlen;
};
void launch_site_amps(const SiteAmpJob* jobs, int njobs, hipStream_t s);

void launch_one_qubit(const OneQubitJob* jobs, int njobs, int max_elems, hipStream_t s);
void launch_merge(const MergeJob* jobs, int njobs, int max_dl, int max_dr, hipStream_t s);
void launch_sumsq(const double2* const* ptrs, const long* counts, double* partial,
                  int njobs, int nblk, hipStream_t s);
void launch_sumsq_final(const double* partial, double* out, int njobs, int nblk, hipStream_t s);
void launch_dust(double2* const* ptrs, const long* counts, const double* fro2, double rel,
                 int njobs, long max_count, hipStream_t s);
void launch_identity(double2* const* ptrs, const int* dims, int njobs, int max_dim,
                     hipStream_t s);
// wide: the 512-thread form, for rounds with at most one block pair per CU.
void launch_jacobi_round(const JacobiJob* jobs, int njobs, int max_pairs, int round,
                         double tol, double nu2, int inner_cap, int* rotated, const int* done,
                         unsigned long long* stats, unsigned long long* offmax, bool wide, hipStream_t s);
void launch_jacobi_round_lo(const JacobiLoJob* jobs, int njobs, int max_pairs, int round,
                            float tol, double nu2, float rel_floor, int inner_cap, int* rotated, const int* done,
                            unsigned long long* stats, unsigned long long* offmax, bool wide, bool bf16_gram,
                            hipStream_t s);
void launch_to_f32(const double2* const* src, float2* const* dst, const long* counts, int njobs,
                   long max_count, hipStream_t s);
void launch_to_f64(const float2* const* src, double2* const* dst, const long* counts, int njobs,
                   long max_count, hipStream_t s);
// jobs[j].dst[0 .. n) = jobs[j].src[0 .. n) in doubles, one launch for the batch.
void launch_copy_batch(const CopyJob* jobs, int njobs, long max_count, hipStream_t s);
// The factorization check of a split: d and r are n x 2 row-major (the two probe columns of
// A X and of the rebuilt factors); out[0..1] receives sum_i |d[i, p] - r[i, p]|^2 per column p.
struct ProbeJob {
    const double2* d;
    const double2* r;
    double* out;
    int n, pad;
};
void launch_probe_residual(const ProbeJob* jobs, int njobs, hipStream_t s);
void launch_row_norms(const JacobiJob* jobs, int njobs, int max_n, double* sigma,
                      const long* sigma_off, hipStream_t s);
// tile: 64 (default) or 32 (four times the workgroups, the same bits).
void launch_gemm(const GemmJob* jobs, int njobs, int max_m, int max_n, hipStream_t s, int tile = 64);
void launch_gemm_splitk(const GemmJob* jobs, int njobs, int max_m, int max_n, int splits, hipStream_t s,
                        int tile = 64);
void launch_factor(const FactorJob* jobs, int njobs, long max_left, long max_right,
                   hipStream_t s);
void launch_householder(const HouseJob* jobs, int njobs, hipStream_t s);
// Panel starting at row j0 of the block starting at row b0, for every job with n > j0;
// max_cols = the largest m - b0.
void launch_lq_panel(const LqJob* jobs, int njobs, int b0, int j0, int max_cols, hipStream_t s);
// T of the block starting at row b0 (rows b0 .. b0 + min(LQ_SB, n - b0) - 1).
void launch_lq_tblock(const LqJob* jobs, int njobs, int b0, hipStream_t s);
void launch_lq_extract(const LqJob* jobs, int njobs, int max_n, hipStream_t s);
void launch_lq_init_m(const LqJob* jobs, int njobs, long max_elems, hipStream_t s);
constexpr int kVecMatChunk = 16;  // rows per workgroup of the left pass
void launch_vecmat(const VecMatJob* job, int out_len, int in_len, int right, hipStream_t s);
// Left passes of njobs independent vector-times-site products in one launch pair (max_out: the
// largest site r, max_in: the largest min(l, vlen)); each job has its own part buffer.
void launch_vecmat_batch(const VecMatJob* jobs, int njobs, int max_out, int max_in, hipStream_t s);
void launch_fill_int(int* p, int value, int count, hipStream_t s);
void launch_conj_transpose(const double2* src, double2* dst, int rows, int cols, hipStream_t s);
void launch_add_identity(double2* C, int rows, int cols, int ldc, hipStream_t s);
void launch_statevector_step(const double2* st, const double2* site, double2* next, int l,
                             int r, long prefix, hipStream_t s);

// Expectation values: the Pauli-transformed copies of a site, B_t = sum_s' conj(P[s',t]) A_s'
// for P = X, Y, Z, and the closing contraction out = sum_i E[i] conj(R[i]) of a transfer-matrix
// chain with a conjugated right environment.
struct CloseJob {
    const double2* E;
    const double2* R;
    double2* out;
    long count;
};
void launch_pauli_site(const double2* A, double2* bx, double2* by, double2* bz, int l, int r,
                       hipStream_t s);
void launch_env_close(const CloseJob* jobs, int njobs, hipStream_t s);

// One site of <bra|ket>: En[a', b'] = sum_{a, b, s} E[a, b] K[a, s, a'] conj(B[b, s, b']) with
// E (kl x bl) and En (kr x br) row-major, K (kl, 2, kr) and B (bl, 2, br) sites; kl * bl and
// kr * br at most 256 (one workgroup).
constexpr int OVERLAP_ENV_MAX = 256;
void launch_overlap_site(const double2* E, const double2* K, const double2* B, double2* En, int kl, int kr,
                         int bl, int br, hipStream_t s);

// MPO x MPS site: out[(a, w), t, (b, v)] = sum_s W[w][v][t][s] A[a, s, b] for MPO bond indices
// w in [w0, w0 + wn) of Dl and v in [v0, v0 + vn) of Dr (W is Dl x Dr x 2 x 2, row-major);
// the combined bond index is mps_index * wn + (w - w0). out is (l*wn) x 2 x (r*vn).
void launch_mpo_site(const double2* A, const double2* W, double2* out, int l, int r, int Dl, int Dr,
                     int w0, int wn, int v0, int vn, hipStream_t s);

// Adjoint sweep. out[2s + t] = sum_{a,b} conj(C[a,s,b]) S[a,t,b] for a site cotangent C and
// a site S, both (l, 2, r).
struct SiteCorrJob {
    const double2* C;
    const double2* S;
    double2* out;
    int l, r;
};
void launch_site_corr(const SiteCorrJob* jobs, int njobs, hipStream_t s);
// The 4x4 gate of a unit as a kernel argument.
struct U4Arg {
    double2 m[16];
};
// T (2dl x 2dr, row-major (a, s, t, b)) <- the cotangent through U4: sum over (s', t') of
// conj(U4[(s', t')][(s, t)]) T[a, s', t', b], in place.
void launch_apply4_adj(double2* T, int dl, int dr, const double2 u4[16], hipStream_t s);
// Coefficients of the adjoint of one truncated split W = U S V^H (W p x q, p <= q; the Jacobi's
// rotated rows X = J W have norms sig, U = J^H; rows krow[0..c) kept, inv[row] = kept index or
// -1). JU = J Abar and XV = X Bbar^H (p x c, unscaled) from the cotangents of the factors
// A = U_k sqrt(S), B = sqrt(S) V_k^H; out: C (p x p) and E (p x c) with
// Wbar = J^H (C X + E Bbar). Rows with sig^2 <= floor2 count as null.
struct SvdAdjJob {
    const double2* JU;
    const double2* XV;
    const double* sig;
    const int* inv;
    double2* C;
    double2* E;
    int p, c;
    double floor2;
};
void launch_svd_adj_coef(const SvdAdjJob& job, hipStream_t s);

// One site of the MPS sampler for `shots` shots: W (shots x 2r) holds v A_s for s = 0, 1,
// Z (two blocks of shots x r) holds R (v A_s)^T; picks s with the uniform u[k * ustride],
// sets bit `bit` of bits[k], writes v' = (v A_s) / sqrt(p_s) into vnext (shots x r).
void launch_sample_step(const double2* W, const double2* Z, int r, long shots, const double* u,
                        long ustride, int bit, unsigned long long* bits, double2* vnext,
                        hipStream_t s);

}  // namespace wqc

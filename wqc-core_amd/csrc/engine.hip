// engine.hip — host side of the MI355X MPS engine and its C ABI (include/wqc_tn.h).
//
// The state lives in device memory: one buffer per site tensor (l, 2, r), sized for the
// largest bond the site can reach (min(chi, 2^i, 2^(n-i)) on each side). Gates are
// scheduled into layers of site-disjoint operations; each layer's one-qubit updates are
// one launch and its two-site units run as batched launches (merge + U, split). The host
// reads back only the singular values of each split (to apply the truncation rule in
// FP64 exactly as the reference does) and the per-sweep Jacobi convergence flags.
//
// Reference items restated here (paths relative to the reference repository root):
//   MpsState::try_new_with_bond       src/tn/mps.rs:82-126
//   apply_gate / apply_one_qubit      src/tn/mps.rs:278-351
//   apply_two_qubit_gate (SWAP route) src/tn/mps.rs:353-375
//   apply_two_qubit_at_positions      src/tn/mps.rs:377-429
//   apply_adjacent_disjoint_gates     src/tn/mps.rs:520-565
//   split_theta / rsvd_sites          src/tn/mps.rs:929-1053
//   truncated_svd_split               src/tn/mps.rs:1124-1201
//   checked_svd dust rule + norm check src/tn/svd_guard.rs:91-139, 242-256
//   retained_chi / needs_f64_fallback src/tn/svd_guard.rs:12-76
//   apply_ccnot                       src/tn/mps.rs:657-677
//   readouts                          src/tn/mps.rs:215-276, 679-740
//   execute_with_trace apply loop     src/tn/trace.rs:71-161

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "../../include/wqc_tn.h"
#include "born.h"
#include "kernels.h"

namespace {

using wqc::FactorJob;
using wqc::GemmJob;
using wqc::HouseJob;
using wqc::JacobiJob;
using wqc::MergeJob;
using wqc::OneQubitJob;
using wqc::VecMatJob;

thread_local std::string g_last_error;

struct Failure {
    int code;
    std::string msg;
};

[[noreturn]] void fail(int code, const std::string& msg) { throw Failure{code, msg}; }

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess)
        fail(e == hipErrorOutOfMemory ? WQC_TN_OUT_OF_MEMORY : WQC_TN_EXEC_FAILED,
             std::string(what) + ": " + hipGetErrorString(e));
}
#define HIPC(x) hip_check((x), #x)

template <class F>
int guarded(F&& f) {
    try {
        f();
        g_last_error.clear();
        return WQC_TN_OK;
    } catch (const Failure& e) {
        g_last_error = e.msg;
        return e.code;
    } catch (const std::bad_alloc&) {
        g_last_error = "host allocation failed";
        return WQC_TN_OUT_OF_MEMORY;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return WQC_TN_EXEC_FAILED;
    }
}

// ---------------------------------------------------------------- constants of the rule
const double SINGULAR_CUTOFF = 1e-12;       // src/tn/svd_guard.rs:12
const double CHI_REL_GAP = 1e-3;            // src/tn/svd_guard.rs:15
const double CUTOFF_ABS_BAND = 1e-10;       // src/tn/svd_guard.rs:18
const double DUST_REL = 1e-14;              // src/tn/svd_guard.rs:93
const double SVD_CHECK_TOL = 1e-9;          // src/tn/svd_guard.rs:234
const size_t RSVD_MIN_SIDE = 160;           // src/tn/mps.rs:794
const size_t RSVD_OVERSAMPLE = 16;          // src/tn/mps.rs:795
// Range-finder rank rule: a column of Y whose Householder residual is at most this
// fraction of its norm is dependent (identity reflector), as in exact arithmetic.
const double QR_RANK_TOL = 1e-10;
const int MAX_SWEEPS = 60;
// Rows whose squared norm is below NULL_ROWS_SQ / n are not rotated: together they carry at
// most 1e-13 of singular value, an order below SINGULAR_CUTOFF.
const double NULL_ROWS_SQ = 1e-26;
// Relative orthogonality tolerances of the Jacobi sweeps, times sqrt(row length): for FP64 an
// order above the rounding of a Gram entry over that many terms; the mixed form's low rounds
// stop at 1e-4 (where the FP64 finish takes over with quadratic convergence; measured best of
// 1e-6, 1e-5, 1e-4 on saturated chi = 512 splits).
const double JACOBI_TOL = 1e-14;
// A job whose largest relative off-diagonal is at most this and no longer halves per sweep
// stops: its rotations have reached their rounding floor (rows orthogonal to 3e-11 leave
// singular values exact to ~1e-21 relative and vectors to ~3e-11). WQC_TN_JACOBI_STALL
// overrides it (0 turns the rule off).
const double STALL_OFF = 3e-11;
const double JACOBI_TOL_LO = 1e-4;
const int MAX_SWEEPS_LO = 30;
// The low rounds run at most this many sweeps per job: they pay while the off-diagonal mass
// falls linearly; after that, rows near FP32 resolution keep single jobs rotating without
// helping the FP64 finish (config 3 with LQ: cap 30 -> 939 merges/s, 6 -> 1387, 3 -> 1363;
// with the current round kernels 8 -> 1537, 7 -> 1561, 6 -> 1588, 5 -> 1602, 4 -> 1601).
const int LO_SWEEP_CAP = 5;
// With LQ preconditioning (the mixed form), Jacobi jobs of at least this many rows (and at
// most LQ_MAX_COLS columns) are rotated as R = L^H of a Householder LQ, X = L Q: on graded
// and rank-deficient spectra the rows of R converge in about half the sweeps of the rows of
// X, and a wide X (m = 2n) leaves an n x n R, half the work per sweep.
const int LQ_MIN_N = 64;

size_t retained_chi(const std::vector<double>& v, size_t max_bond, size_t rows, size_t cols) {
    if (v.empty() || max_bond == 0 || rows == 0 || cols == 0) return 0;
    size_t rank = 0;
    for (double s : v)
        if (s > SINGULAR_CUTOFF) rank++;
    rank = std::max<size_t>(rank, 1);
    return std::min(std::min(rank, max_bond), std::min(rows, cols));
}

bool needs_f64_fallback(const std::vector<double>& v, size_t chi) {
    for (double s : v)
        if (!std::isfinite(s) || s < 0.0) return true;
    for (double s : v)
        if (s > 0.0 && std::fabs(s - SINGULAR_CUTOFF) <= CUTOFF_ABS_BAND) return true;
    if (chi > 0 && chi < v.size()) {
        double kept = v[chi - 1], dropped = v[chi];
        double scale = std::max(std::max(kept, dropped), 1e-30);
        if (std::fabs(kept - dropped) / scale < CHI_REL_GAP) return true;
    }
    return false;
}

bool valid_precision(int p) {
    return p == WQC_TN_PRECISION_FP64 || p == WQC_TN_PRECISION_MIXED_FP32 || p == WQC_TN_PRECISION_MIXED_BF16;
}

bool rsvd_worth(size_t rows, size_t cols, size_t ell) {
    unsigned __int128 f = (unsigned __int128)rows * cols * ell;
    unsigned __int128 side = RSVD_MIN_SIDE;
    return f >= side * side * side;
}

// ---------------------------------------------------------------- gate tensors
typedef double2 Mat2[4];    // row-major 2x2
typedef double2 Mat4[16];   // [c_out][t_out][c_in][t_in]

inline double2 cpx(double r, double i) { return make_double2(r, i); }

bool is_one_qubit(int k) {
    return k == WQC_TN_GATE_H || k == WQC_TN_GATE_X || k == WQC_TN_GATE_Y || k == WQC_TN_GATE_Z ||
           k == WQC_TN_GATE_S || k == WQC_TN_GATE_T || k == WQC_TN_GATE_RX ||
           k == WQC_TN_GATE_RY || k == WQC_TN_GATE_RZ;
}

// unary_matrix (src/tn/gates.rs:10-71); kind 0 = T-dagger (src/tn/mps.rs:1173-1180).
void unary_matrix(int kind, double angle, Mat2 u) {
    const double fs2 = 0.70710678118654752440;
    for (int i = 0; i < 4; ++i) u[i] = cpx(0, 0);
    switch (kind) {
        case WQC_TN_GATE_H: {
            double k = 1.0 / std::sqrt(2.0);
            u[0] = cpx(k, 0);
            u[1] = cpx(k, 0);
            u[2] = cpx(k, 0);
            u[3] = cpx(-k, -0.0);
            break;
        }
        case WQC_TN_GATE_X:
            u[1] = cpx(1, 0);
            u[2] = cpx(1, 0);
            break;
        case WQC_TN_GATE_Y:
            u[1] = cpx(-0.0, -1.0);
            u[2] = cpx(0.0, 1.0);
            break;
        case WQC_TN_GATE_Z:
            u[0] = cpx(1, 0);
            u[3] = cpx(-1.0, -0.0);
            break;
        case WQC_TN_GATE_S:
            u[0] = cpx(1, 0);
            u[3] = cpx(0, 1);
            break;
        case WQC_TN_GATE_T:
            u[0] = cpx(1, 0);
            u[3] = cpx(fs2, fs2);
            break;
        case WQC_TN_GATE_RX: {
            double s = std::sin(angle / 2.0), c = std::cos(angle / 2.0);
            u[0] = cpx(c, 0);
            u[1] = cpx(0, -s);
            u[2] = cpx(0, -s);
            u[3] = cpx(c, 0);
            break;
        }
        case WQC_TN_GATE_RY: {
            double s = std::sin(angle / 2.0), c = std::cos(angle / 2.0);
            u[0] = cpx(c, 0);
            u[1] = cpx(-s, 0);
            u[2] = cpx(s, 0);
            u[3] = cpx(c, 0);
            break;
        }
        case WQC_TN_GATE_RZ: {
            double s = std::sin(angle / 2.0), c = std::cos(angle / 2.0);
            u[0] = cpx(c, -s);
            u[3] = cpx(c, s);
            break;
        }
        case 0:
            u[0] = cpx(1, 0);
            u[3] = cpx(fs2, -fs2);
            break;
        default:
            fail(WQC_TN_INVALID_ARG, "not a one-qubit gate");
    }
}

inline int m4(int co, int to, int ci, int ti) { return ((co * 2 + to) * 2 + ci) * 2 + ti; }

// cnot_matrix / cz_matrix / swap_matrix (src/tn/gates.rs:73-104).
void two_qubit_matrix(int kind, Mat4 u) {
    for (int i = 0; i < 16; ++i) u[i] = cpx(0, 0);
    if (kind == WQC_TN_GATE_CNOT) {
        u[m4(0, 0, 0, 0)] = cpx(1, 0);
        u[m4(0, 1, 0, 1)] = cpx(1, 0);
        u[m4(1, 0, 1, 1)] = cpx(1, 0);
        u[m4(1, 1, 1, 0)] = cpx(1, 0);
    } else if (kind == WQC_TN_GATE_CZ) {
        u[m4(0, 0, 0, 0)] = cpx(1, 0);
        u[m4(0, 1, 0, 1)] = cpx(1, 0);
        u[m4(1, 0, 1, 0)] = cpx(1, 0);
        u[m4(1, 1, 1, 1)] = cpx(-1.0, -0.0);
    } else {  // SWAP
        u[m4(0, 0, 0, 0)] = cpx(1, 0);
        u[m4(0, 1, 1, 0)] = cpx(1, 0);
        u[m4(1, 0, 0, 1)] = cpx(1, 0);
        u[m4(1, 1, 1, 1)] = cpx(1, 0);
    }
}
const int KIND_SWAP = 100;

// u4[(so*2+to)*4 + si*2+ti]: the unitary on the (left site, right site) physical pair
// (the ctrl_is_first index mapping of src/tn/mps.rs:404-414).
void resolve_u4(const Mat4 u, bool ctrl_is_first, double2* u4) {
    for (int so = 0; so < 2; ++so)
        for (int to = 0; to < 2; ++to)
            for (int si = 0; si < 2; ++si)
                for (int ti = 0; ti < 2; ++ti) {
                    int co = ctrl_is_first ? so : to, tco = ctrl_is_first ? to : so;
                    int ci = ctrl_is_first ? si : ti, tci = ctrl_is_first ? ti : si;
                    u4[(so * 2 + to) * 4 + si * 2 + ti] = u[m4(co, tco, ci, tci)];
                }
}

// ---------------------------------------------------------------- device buffers
// Grow-only device buffer; freed by release() or when its owner goes away (the owner makes
// its device current first). Not copyable: one owner per allocation.
// Device blocks that a device's states have let go, kept for the next buffer of about that
// size on the same device. A state grows its workspaces as the bonds grow, callers typically
// run one circuit per state, and hipFree waits for the whole device: without the cache every
// circuit pays a few dozen reallocations and a hundred frees, each with the device drained.
// A block carries an event, recorded on the stream of its last user when that user may still
// have work queued on it; the stream that takes the block waits for that event on the device.
// The cache holds at most kHeldMax bytes; it is emptied when a hipMalloc runs out of memory
// (and the allocation tried again) and when the device handle is closed.
struct BlockPool {
    struct Block {
        void* p = nullptr;
        size_t bytes = 0;
        hipEvent_t ev = nullptr;
    };
    static constexpr size_t kHeldMax = size_t(48) << 30;
    std::mutex mu;
    std::multimap<size_t, Block> blocks;
    size_t held = 0;

    static size_t round_up(size_t want) {
        const size_t g = want < (size_t(1) << 20) ? size_t(64) << 10 : size_t(1) << 20;
        return (std::max<size_t>(want, 1) + g - 1) / g * g;
    }
    // A block of at least `want` bytes for work queued on stream s: a cached one of at most
    // twice that (plus 1 MiB), else a new one.
    Block get(size_t want, hipStream_t s) {
        want = round_up(want);
        Block b;
        bool hit = false;
        {
            std::lock_guard<std::mutex> lk(mu);
            auto it = blocks.lower_bound(want);
            if (it != blocks.end() && it->first <= 2 * want + (size_t(1) << 20)) {
                b = it->second;
                held -= b.bytes;
                blocks.erase(it);
                hit = true;
            }
        }
        if (hit) {
            if (b.ev) {
                const hipError_t e = s ? hipStreamWaitEvent(s, b.ev, 0) : hipEventSynchronize(b.ev);
                if (e != hipSuccess) {
                    put(b, nullptr);
                    hip_check(e, "waiting for a cached block");
                }
            }
            return b;
        }
        b.bytes = want;
        hipError_t e = hipMalloc(&b.p, want);
        if (e == hipErrorOutOfMemory) {
            (void)hipGetLastError();
            release_all();
            e = hipMalloc(&b.p, want);
        }
        hip_check(e, "hipMalloc");
        return b;
    }
    // Returns a block; s: the stream whose queued work may still use it (null: none may).
    void put(Block b, hipStream_t s) noexcept {
        if (!b.p) return;
        if (s) {
            bool ok = b.ev || hipEventCreateWithFlags(&b.ev, hipEventDisableTiming) == hipSuccess;
            ok = ok && hipEventRecord(b.ev, s) == hipSuccess;
            if (!ok) {
                (void)hipStreamSynchronize(s);
                (void)hipFree(b.p);
                if (b.ev) (void)hipEventDestroy(b.ev);
                return;
            }
        }
        bool kept = true;
        std::multimap<size_t, Block> drop;  // filled by node moves, which allocate nothing
        {
            std::lock_guard<std::mutex> lk(mu);
            try {
                blocks.emplace(b.bytes, b);
                held += b.bytes;
            } catch (...) {
                kept = false;  // no host memory for the entry: the block is freed below
            }
            while (held > kHeldMax && !blocks.empty()) {
                auto node = blocks.extract(std::prev(blocks.end()));
                held -= node.key();
                drop.insert(std::move(node));
            }
        }
        if (!kept) {
            if (s) (void)hipStreamSynchronize(s);
            (void)hipFree(b.p);
            if (b.ev) (void)hipEventDestroy(b.ev);
        }
        for (auto& kv : drop) {
            (void)hipFree(kv.second.p);
            if (kv.second.ev) (void)hipEventDestroy(kv.second.ev);
        }
    }
    size_t held_bytes() {
        std::lock_guard<std::mutex> lk(mu);
        return held;
    }
    void release_all() noexcept {
        std::multimap<size_t, Block> all;
        {
            std::lock_guard<std::mutex> lk(mu);
            all.swap(blocks);
            held = 0;
        }
        for (auto& kv : all) {
            (void)hipFree(kv.second.p);
            if (kv.second.ev) (void)hipEventDestroy(kv.second.ev);
        }
    }
    ~BlockPool() { release_all(); }
};

// A device buffer that only grows (its contents are not kept when it does). Bound to a pool
// and a stream, it takes its blocks from the device's cache and returns them there, growing by
// at least half its size each time; unbound, it calls hipMalloc and hipFree itself.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    BlockPool* pool = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t ev = nullptr;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void bind(BlockPool* bp, hipStream_t s) {
        release();
        pool = bp;
        stream = s;
    }
    void ensure(size_t want) {
        if (want <= bytes) return;
        if (pool) {
            const size_t sz = std::max(want, bytes + bytes / 2);
            release();
            const BlockPool::Block b = pool->get(sz, stream);
            p = b.p;
            bytes = b.bytes;
            ev = b.ev;
            return;
        }
        if (p) HIPC(hipFree(p));
        p = nullptr;
        bytes = 0;
        size_t sz = std::max(want, (size_t)1 << 20);
        sz = (sz + ((size_t)1 << 20) - 1) & ~(((size_t)1 << 20) - 1);
        HIPC(hipMalloc(&p, sz));
        bytes = sz;
    }
    // synced: no work queued on the stream uses the buffer any more.
    void release(bool synced = false) {
        if (pool) {
            if (p) pool->put(BlockPool::Block{p, bytes, ev}, synced ? nullptr : stream);
        } else if (p) {
            (void)hipFree(p);
        }
        p = nullptr;
        bytes = 0;
        ev = nullptr;
    }
};

// Bump allocator over device chunks of at least 256 MiB, freed together (the adjoint sweep's
// per-segment record of a forward replay).
struct TapeArena {
    std::vector<std::unique_ptr<DevBuf>> chunks;
    size_t used = 0;
    double2* take(size_t elems) {
        const size_t bytes = ((elems * sizeof(double2)) + 255) & ~(size_t)255;
        if (chunks.empty() || used + bytes > chunks.back()->bytes) {
            chunks.push_back(std::make_unique<DevBuf>());
            chunks.back()->ensure(std::max(bytes, (size_t)256 << 20));
            used = 0;
        }
        double2* p = reinterpret_cast<double2*>((char*)chunks.back()->p + used);
        used += bytes;
        return p;
    }
};

// Forward replay recorded for the adjoint sweep, in execution order: every one-qubit update
// (site, matrix, the gate it came from) and every two-site unit with its input sites, the
// Jacobi's rotated rows X = J W (W = theta' or theta'^H, p x q, p <= q), the rotation J (p x p),
// the row norms (Jacobi order) and the kept rows in order.
struct TapeOne {
    int site, tag;
    double2 u[4];
};
struct TapeUnit {
    int p = 0, dl = 0, bond = 0, dr = 0, pp = 0, qq = 0, c = 0, tag = -1;
    bool tall = false, done = false;
    double2 u4[16];
    std::vector<int> krow;
    std::vector<double> sig;
    double2 *L = nullptr, *R = nullptr, *X = nullptr, *J = nullptr;
};
struct Tape {
    std::vector<std::pair<int, int>> order;  // (0, index into ones) or (1, index into units)
    std::vector<TapeOne> ones;
    std::vector<TapeUnit> units;
    TapeArena arena;
    int tag = -1;  // the gate being replayed
};

struct OmegaEntry {
    double2* omega_cols = nullptr;  // cols x r row-major: the independent columns of Omega
    int* pos = nullptr;             // r positions
    int r = 0;
};

}  // namespace

// ---------------------------------------------------------------- allocation-free scratch
// Device memory of wqc_tn_amplitude_at_compact_zero_noalloc: its two chain vectors and the left
// pass's row-chunk partials. Every GPU holds its own copy for the life of the process; host code
// never uses this name as an address, only what hipGetSymbolAddress returns for the current
// device (kept per device in wqc_tn_device::scratch). Calls that overlap on one GPU race on it.
constexpr size_t kScratchElems = size_t(1) << 17;  // 2 MiB of complex128
static __device__ __align__(256) double2 scratch[kScratchElems];

// ---------------------------------------------------------------- device handle
struct wqc_tn_device {
    int ordinal = 0;
    hipDeviceProp_t prop{};
    double2* scratch = nullptr;  // this device's address of the scratch array
    std::mutex mu;
    std::map<std::tuple<size_t, size_t, size_t>, OmegaEntry> omega;
    // GEMM autotuner: the tile (64 or 32) measured fastest per shape class (log2 buckets of the
    // largest M, N, K and of the job count, split-K factor, operand ops), kept for the process
    // and, with WQC_TN_TUNE_CACHE=<file>, in that file across processes.
    std::map<std::array<int, 6>, int> gemm_tile;
    std::string tune_path;
    // Device blocks and pinned host staging that closed states have let go, for the next state.
    BlockPool pool;
    std::mutex pin_mu;
    std::vector<std::pair<void*, size_t>> pinned;
    void* get_pinned(size_t bytes) {
        {
            std::lock_guard<std::mutex> lk(pin_mu);
            for (size_t i = 0; i < pinned.size(); ++i)
                if (pinned[i].second == bytes) {
                    void* p = pinned[i].first;
                    pinned.erase(pinned.begin() + i);
                    return p;
                }
        }
        void* p = nullptr;
        HIPC(hipHostMalloc(&p, bytes, hipHostMallocDefault));
        return p;
    }
    void put_pinned(void* p, size_t bytes) noexcept {  // no queued copy may still use p
        if (!p) return;
        std::lock_guard<std::mutex> lk(pin_mu);
        if (pinned.size() < 16) {
            pinned.emplace_back(p, bytes);
            return;
        }
        (void)hipHostFree(p);
    }
    void load_tune_cache() {
        if (tune_path.empty()) return;
        FILE* f = std::fopen(tune_path.c_str(), "r");
        if (!f) return;
        std::array<int, 6> k{};
        int tile = 0;
        while (std::fscanf(f, "gemm %d %d %d %d %d %d %d\n", &k[0], &k[1], &k[2], &k[3], &k[4], &k[5], &tile) == 7)
            if (tile == 32 || tile == 64) gemm_tile[k] = tile;
        std::fclose(f);
    }
    void save_tune_entry(const std::array<int, 6>& k, int tile) {  // caller holds mu
        if (tune_path.empty()) return;
        if (FILE* f = std::fopen(tune_path.c_str(), "a")) {
            std::fprintf(f, "gemm %d %d %d %d %d %d %d\n", k[0], k[1], k[2], k[3], k[4], k[5], tile);
            std::fclose(f);
        }
    }
    ~wqc_tn_device() {
        (void)hipSetDevice(ordinal);
        for (auto& pp : pinned) (void)hipHostFree(pp.first);
        pool.release_all();
        for (auto& kv : omega) {
            if (kv.second.omega_cols) (void)hipFree(kv.second.omega_cols);
            if (kv.second.pos) (void)hipFree(kv.second.pos);
        }
    }
};

namespace {

// rademacher / mix_seed (src/tn/mps.rs:1103-1122), then the independent columns of
// Omega in order (modified Gram-Schmidt with one re-orthogonalization, relative residual
// threshold QR_RANK_TOL). A column of Omega that depends on earlier columns gives a
// column of Y = A A^H A Omega that depends on the same earlier columns, so only these
// columns can carry a Householder reflector.
OmegaEntry& omega_for(wqc_tn_device* dev, size_t rows, size_t cols, size_t ell) {
    std::lock_guard<std::mutex> lk(dev->mu);
    auto key = std::make_tuple(rows, cols, ell);
    auto it = dev->omega.find(key);
    if (it != dev->omega.end()) return it->second;
    uint64_t salt = ((uint64_t)rows * 0x9E37ull + (uint64_t)cols) * 0x9E37ull + (uint64_t)ell;
    uint64_t state = salt | 1ull;
    const double scale = 1.0 / std::sqrt((double)cols);
    std::vector<double2> om(cols * ell);
    for (size_t i = 0; i < cols * ell; ++i) {
        state = state * 6364136223846793005ull + 1ull;
        om[i] = make_double2((state & 0x100) == 0 ? scale : -scale, (state & 0x200) == 0 ? scale : -scale);
    }
    std::vector<std::vector<double2>> basis;
    std::vector<int> pos;
    std::vector<double2> w(cols);
    for (size_t e = 0; e < ell; ++e) {
        double n0 = 0.0;
        for (size_t c = 0; c < cols; ++c) {
            w[c] = om[c * ell + e];
            n0 += w[c].x * w[c].x + w[c].y * w[c].y;
        }
        n0 = std::sqrt(n0);
        for (int pass = 0; pass < 2; ++pass)
            for (auto& b : basis) {
                double dr = 0, di = 0;  // b^H w
                for (size_t c = 0; c < cols; ++c) {
                    dr += b[c].x * w[c].x + b[c].y * w[c].y;
                    di += b[c].x * w[c].y - b[c].y * w[c].x;
                }
                for (size_t c = 0; c < cols; ++c) {
                    w[c].x -= b[c].x * dr - b[c].y * di;
                    w[c].y -= b[c].x * di + b[c].y * dr;
                }
            }
        double nw = 0.0;
        for (size_t c = 0; c < cols; ++c) nw += w[c].x * w[c].x + w[c].y * w[c].y;
        nw = std::sqrt(nw);
        if (nw > QR_RANK_TOL * n0) {
            std::vector<double2> b(cols);
            for (size_t c = 0; c < cols; ++c) b[c] = make_double2(w[c].x / nw, w[c].y / nw);
            basis.push_back(std::move(b));
            pos.push_back((int)e);
        }
        if (basis.size() >= std::min(cols, ell)) break;
    }
    OmegaEntry ent;
    ent.r = (int)pos.size();
    std::vector<double2> sel(cols * ent.r);
    for (size_t c = 0; c < cols; ++c)
        for (int k = 0; k < ent.r; ++k) sel[c * ent.r + k] = om[c * ell + pos[k]];
    HIPC(hipMalloc(&ent.omega_cols, sel.size() * sizeof(double2)));
    HIPC(hipMemcpy(ent.omega_cols, sel.data(), sel.size() * sizeof(double2), hipMemcpyHostToDevice));
    HIPC(hipMalloc(&ent.pos, std::max<size_t>(1, pos.size()) * sizeof(int)));
    HIPC(hipMemcpy(ent.pos, pos.data(), pos.size() * sizeof(int), hipMemcpyHostToDevice));
    return dev->omega.emplace(key, ent).first->second;
}

struct Site {
    double2* ptr = nullptr;
    size_t cap = 0;  // complex elements
    int l = 1, r = 1;
    bool owned = true;  // false: carved from the state's site arena (freed with it)
    size_t block = 0;   // owned and from the device's block cache: the block's bytes
    hipEvent_t ev = nullptr;  // that block's event
};

// One two-site unit: sites p and p+1, resolved 4x4 unitary.
}  // namespace

struct wqc_tn_state;

namespace {
// One two-site unit: sites p and p+1 of `owner` (null: the state that runs it), resolved 4x4
// unitary. A batch of independent slices runs units of many states in one launch.
struct Unit {
    int p;
    double2 u4[16];
    wqc_tn_state* owner = nullptr;
};

struct PendingOne {
    int site;
    double2 u[4];
    wqc_tn_state* owner = nullptr;
};

}  // namespace

// ---------------------------------------------------------------- state
struct wqc_tn_state {
    wqc_tn_device* dev = nullptr;
    int n = 0;
    size_t chi = 0;
    bool reject = false;
    bool exact_only = false;  // every split exact (the adjoint state of the gradient sweep)
    int precision = WQC_TN_PRECISION_FP64;
    std::vector<Site> sites;
    hipStream_t stream = nullptr;
    // Streams for the further groups of a Jacobi batch's jobs (group 0 runs on `stream`),
    // forked from and joined back into `stream` by ev_fork / ev_join; ev_sweep[g] marks group
    // g's latest sweep.
    static constexpr int kMaxGroups = 4;
    hipStream_t streams_x[kMaxGroups - 1] = {};
    hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_sweep[kMaxGroups] = {};
    DevBuf ws;        // split workspace
    DevBuf ws32;      // mixed-precision rows and the J X product
    DevBuf ews;       // expectation environments and transfer buffers
    DevBuf tab;       // device job tables
    DevBuf gather;    // sigma and |A|_F^2 of a Jacobi batch, gathered for one readback
    char* htab = nullptr;  // pinned staging for tables
    size_t htab_cap = 0, tab_off = 0;
    double* hread = nullptr;  // pinned readback
    size_t hread_cap = 0;
    std::atomic<int> cancel{0};
    wqc_tn_stats stats{};
    uint64_t site_bytes = 0;
    bool timing = false;
    double kms[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    uint64_t klaunch[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    struct TimedSpan {
        hipEvent_t a, b;
        int fam;
    };
    std::vector<TimedSpan> spans;
    std::vector<hipEvent_t> event_pool;
    bool split_log_on = false;
    std::vector<int32_t> split_log;
    std::vector<double> split_dw;  // discarded weight of each logged unit
    std::vector<double> split_probe;  // relative probe residual of each logged unit
    std::vector<PendingOne> pend1;
    std::vector<Unit> pend2;
    std::vector<char> touched;
    Tape* tape = nullptr;  // set while the adjoint sweep replays a segment

    ~wqc_tn_state() {
        (void)hipSetDevice(dev ? dev->ordinal : 0);
        if (stream) (void)hipStreamSynchronize(stream);
        for (hipStream_t x : streams_x)
            if (x) (void)hipStreamSynchronize(x);
        for (auto& s : sites) free_site(s, true);
        for (DevBuf* b : state_bufs()) b->release(true);
        for (auto& sp : spans) {
            (void)hipEventDestroy(sp.a);
            (void)hipEventDestroy(sp.b);
        }
        for (auto e : event_pool) (void)hipEventDestroy(e);
        if (dev) {
            dev->put_pinned(htab, htab_cap);
            dev->put_pinned(hread, hread_cap * sizeof(double));
        } else {
            if (htab) (void)hipHostFree(htab);
            if (hread) (void)hipHostFree(hread);
        }
        for (hipEvent_t e : {ev_fork, ev_join})
            if (e) (void)hipEventDestroy(e);
        for (hipEvent_t e : ev_sweep)
            if (e) (void)hipEventDestroy(e);
        for (hipStream_t x : streams_x)
            if (x) (void)hipStreamDestroy(x);
        if (stream) (void)hipStreamDestroy(stream);
    }

    // The state's own device buffers (bound to the device's block cache when the state has a
    // stream).
    std::vector<DevBuf*> state_bufs() {
        return {&site_arena, &ws, &ws32, &ews, &tab, &jstats_buf, &trace_buf, &splitk_buf,
                &gather, &ov_buf, &env_buf, &amp_buf, &probe_buf, &probe_p_buf};
    }
    // Gives an owned site's memory back (synced: no queued work reads it any more).
    void free_site(Site& s, bool synced) {
        if (s.ptr && s.owned) {
            if (s.block)
                dev->pool.put(BlockPool::Block{s.ptr, s.block, s.ev}, synced ? nullptr : stream);
            else
                (void)hipFree(s.ptr);
        }
        s.ptr = nullptr;
        s.block = 0;
        s.ev = nullptr;
    }

    // ------------------------------------------------------------ helpers
    // Every device buffer the state owns (sites, workspaces, staging, readback and trace
    // buffers), so peak_vram_bytes is the largest footprint the state has had.
    void note_peak() {
        uint64_t now = site_bytes + ws.bytes + ws32.bytes + ews.bytes + tab.bytes + splitk_buf.bytes + gather.bytes +
                       jstats_buf.bytes + trace_buf.bytes + ov_buf.bytes + env_buf.bytes + probe_buf.bytes +
                       probe_p_buf.bytes + amp_buf.bytes;
        if (now > stats.peak_vram_bytes) stats.peak_vram_bytes = now;
    }

    size_t bond_cap(int i) const {  // bond between sites i-1 and i
        size_t b = chi;
        int e = std::min(i, n - i);
        if (e < 40) b = std::min(b, (size_t)1 << e);
        return std::max<size_t>(b, 1);
    }

    void ensure_site_cap(int i, size_t elems) {
        Site& s = sites[i];
        if (elems <= s.cap) return;
        HIPC(hipStreamSynchronize(stream));
        double2* np = nullptr;
        BlockPool::Block b;
        if (stream) {
            b = dev->pool.get(elems * sizeof(double2), stream);
            np = static_cast<double2*>(b.p);
        } else {
            HIPC(hipMalloc(&np, elems * sizeof(double2)));
        }
        if (s.ptr) {
            if (stream)
                HIPC(hipMemcpyAsync(np, s.ptr, std::min(s.cap, elems) * sizeof(double2), hipMemcpyDeviceToDevice,
                                    stream));
            else
                HIPC(hipMemcpy(np, s.ptr, std::min(s.cap, elems) * sizeof(double2), hipMemcpyDeviceToDevice));
            free_site(s, false);
        }
        site_bytes += s.owned ? (elems - s.cap) * sizeof(double2) : elems * sizeof(double2);
        s.ptr = np;
        s.cap = elems;
        s.owned = true;
        s.block = b.bytes;
        s.ev = b.ev;
        // Callers write the site right away, some through the null stream (which does not
        // wait for this stream): the copy above and a wait for the block's previous user end
        // here.
        if (stream) HIPC(hipStreamSynchronize(stream));
        note_peak();
    }

    // ensure_site_cap on the state that owns the site (this one, or a slice whose unit this
    // state runs): a reallocation of another state's site first drains this state's stream,
    // where the kernels reading that site were queued.
    void grow_site(wqc_tn_state* owner, int i, size_t elems) {
        if (owner == this) {
            ensure_site_cap(i, elems);
            return;
        }
        if (elems <= owner->sites[i].cap) return;
        sync();
        owner->ensure_site_cap(i, elems);
        note_peak();
    }

    void tables_reset() { tab_off = 0; }

    template <class T>
    T* upload(const std::vector<T>& v) {
        size_t bytes = std::max<size_t>(v.size() * sizeof(T), 16);
        size_t off = (tab_off + 255) & ~(size_t)255;
        if (off + bytes > htab_cap) fail(WQC_TN_EXEC_FAILED, "job table arena exhausted");
        std::memcpy(htab + off, v.data(), v.size() * sizeof(T));
        HIPC(hipMemcpyAsync((char*)tab.p + off, htab + off, bytes, hipMemcpyHostToDevice, stream));
        tab_off = off + bytes;
        return reinterpret_cast<T*>((char*)tab.p + off);
    }

    void check_launch() { HIPC(hipGetLastError()); }

    void sync() { HIPC(hipStreamSynchronize(stream)); }

    void check_cancel() {
        if (cancel.load()) fail(WQC_TN_CANCELLED, "cancelled");
    }

    // ------------------------------------------------------------ lifetime
    // Sweeps of the in-block eigensolve per round (WQC_TN_JACOBI_INNER overrides, 1..32).
    int inner_cap = 2;
    int inner_cap_lo = 1;    // the same for the mixed form's low rounds (WQC_TN_JACOBI_INNER_LO)
    double tol_lo = JACOBI_TOL_LO;  // their relative tolerance (WQC_TN_JACOBI_TOL_LO)
    double floor_lo = 1e-12;        // squared-norm floor relative to a pair's largest row (WQC_TN_JACOBI_FLOOR_LO)
    int max_sweeps_lo = LO_SWEEP_CAP;  // cap on the low sweeps per job (WQC_TN_JACOBI_MAX_LO)
    bool trace_jacobi = false;
    double stall_off = STALL_OFF;
    bool use_lq = true;
    bool use_lq2 = true;  // a second factorization R = L2 Q2 before the Jacobi (with use_lq)
    bool use_splitk = true;  // WQC_TN_SPLITK=0 turns split-K GEMMs off
    bool use_wide = true;    // WQC_TN_JACOBI_WIDE=0 keeps every Jacobi round at 256 threads
    DevBuf trace_buf;

    void init(wqc_tn_device* d, size_t qn, size_t mb, int prec) {
        dev = d;
        if (const char* e = std::getenv("WQC_TN_JACOBI_INNER")) {
            const int v = std::atoi(e);
            if (v >= 1 && v <= 32) inner_cap = v;
        }
        if (const char* e = std::getenv("WQC_TN_JACOBI_TRACE")) trace_jacobi = std::atoi(e) != 0;
        if (const char* e = std::getenv("WQC_TN_JACOBI_STALL")) stall_off = std::max(0.0, std::atof(e));
        if (const char* e = std::getenv("WQC_TN_JACOBI_GROUPS")) jacobi_groups = std::min(kMaxGroups, std::max(1, std::atoi(e)));
        if (const char* e = std::getenv("WQC_TN_JACOBI_GROUPS_LO"))
            jacobi_groups_lo = std::min(kMaxGroups, std::max(1, std::atoi(e)));
        // LQ preconditioning: on in the mixed form, off in the plain FP64 form (whose Jacobi on
        // the rows of theta' keeps small singular values relatively accurate; R = L^H carries
        // the LQ's rounding at the scale of the largest rows, which moves near-cutoff
        // decisions); WQC_TN_LQ=0/1 overrides either.
        use_lq = prec != WQC_TN_PRECISION_FP64;
        if (const char* e = std::getenv("WQC_TN_LQ")) use_lq = std::atoi(e) != 0;
        use_lq2 = use_lq;
        if (const char* e = std::getenv("WQC_TN_LQ2")) use_lq2 = use_lq && std::atoi(e) != 0;
        if (const char* e = std::getenv("WQC_TN_SPLITK")) use_splitk = std::atoi(e) != 0;
        if (const char* e = std::getenv("WQC_TN_JACOBI_WIDE")) use_wide = std::atoi(e) != 0;
        if (const char* e = std::getenv("WQC_TN_GRAPH")) use_graph = std::atoi(e) != 0;
        if (const char* e = std::getenv("WQC_TN_AUTOTUNE")) autotune = std::atoi(e) != 0;
        if (const char* e = std::getenv("WQC_TN_GEMM_TILE")) {
            const int t = std::atoi(e);
            if (t == 32 || t == 64) forced_tile = t;
        }
        if (const char* e = std::getenv("WQC_TN_JACOBI_INNER_LO")) {
            const int v = std::atoi(e);
            if (v >= 1 && v <= 32) inner_cap_lo = v;
        }
        if (const char* e = std::getenv("WQC_TN_JACOBI_FLOOR_LO")) {
            const double v = std::atof(e);
            if (v >= 1e-14 && v <= 1e-2) floor_lo = v;
        }
        if (const char* e = std::getenv("WQC_TN_JACOBI_MAX_LO")) {
            const int v = std::atoi(e);
            if (v >= 1 && v <= MAX_SWEEPS_LO) max_sweeps_lo = v;
        }
        if (const char* e = std::getenv("WQC_TN_JACOBI_TOL_LO")) {
            const double v = std::atof(e);
            if (v >= 1e-7 && v <= 1e-2) tol_lo = v;
        }
        n = (int)qn;
        chi = mb;
        precision = prec;
        HIPC(hipSetDevice(dev->ordinal));
        HIPC(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        const int ngroups = std::max(jacobi_groups, jacobi_groups_lo);
        for (int g = 0; g + 1 < ngroups; ++g) HIPC(hipStreamCreateWithFlags(&streams_x[g], hipStreamNonBlocking));
        for (hipEvent_t* e : {&ev_fork, &ev_join}) HIPC(hipEventCreateWithFlags(e, hipEventDisableTiming));
        for (int g = 0; g < ngroups; ++g) HIPC(hipEventCreateWithFlags(&ev_sweep[g], hipEventDisableTiming));
        for (DevBuf* b : state_bufs()) b->bind(&dev->pool, stream);
        htab_cap = 8u << 20;
        htab = static_cast<char*>(dev->get_pinned(htab_cap));
        tab.ensure(htab_cap);
        hread_cap = 1u << 20;
        hread = static_cast<double*>(dev->get_pinned(hread_cap * sizeof(double)));
        sites.assign(n, Site{});
        std::vector<double2> init = {make_double2(1, 0), make_double2(0, 0)};
        for (int i = 0; i < n; ++i) {
            size_t cap = bond_cap(i) * 2 * bond_cap(i + 1);
            ensure_site_cap(i, cap);
            HIPC(hipMemcpy(sites[i].ptr, init.data(), 2 * sizeof(double2), hipMemcpyHostToDevice));
            sites[i].l = sites[i].r = 1;
        }
        touched.assign(n, 0);
        note_peak();
    }

    // A site container for the slice batch: the |0...0> register of qn qubits with bond cap mb,
    // its sites carved from one device arena (light-cone capacities, one allocation), no stream,
    // staging or workspace of its own: another state runs its operations. one_src holds the
    // complex 1 that starts every site (set by the caller's batched copy, see run_slices).
    DevBuf site_arena;
    void init_container(wqc_tn_device* d, size_t qn, size_t mb, int prec, std::vector<wqc::CopyJob>& starts,
                        const double* one_src) {
        dev = d;
        n = (int)qn;
        chi = mb;
        precision = prec;
        sites.assign(n, Site{});
        size_t total = 0;
        std::vector<size_t> at(n);
        for (int i = 0; i < n; ++i) {
            at[i] = total;
            const size_t cap = bond_cap(i) * 2 * bond_cap(i + 1);
            total += (cap + 15) & ~(size_t)15;  // 256-byte aligned sites
        }
        site_arena.ensure(total * sizeof(double2));
        HIPC(hipMemset(site_arena.p, 0, site_arena.bytes));
        double2* base = reinterpret_cast<double2*>(site_arena.p);
        for (int i = 0; i < n; ++i) {
            Site& s = sites[i];
            s.ptr = base + at[i];
            s.cap = bond_cap(i) * 2 * bond_cap(i + 1);
            s.l = s.r = 1;
            s.owned = false;
            starts.push_back(wqc::CopyJob{one_src, reinterpret_cast<double*>(s.ptr), 2, 0});
        }
        site_bytes = site_arena.bytes;
        touched.assign(n, 0);
        note_peak();
    }

    // A fresh |0...0> register of qn qubits and bond cap mb on the same stream, staging and
    // workspaces; site buffers are kept where they are large enough. Used by the slice batch.
    void reset(size_t qn, size_t mb) {
        sync();
        pend1.clear();
        pend2.clear();
        split_log.clear();
        split_dw.clear();
        split_probe.clear();
        stats = wqc_tn_stats{};
        cancel.store(0);
        reject = false;
        std::vector<Site> old = std::move(sites);
        n = (int)qn;
        chi = mb;
        sites.assign(n, Site{});
        for (int i = 0; i < n && i < (int)old.size(); ++i) {
            sites[i] = old[i];
            old[i] = Site{};
        }
        for (auto& s : old)
            if (s.ptr && s.owned) {
                site_bytes -= s.cap * sizeof(double2);
                free_site(s, true);
            }
        static const double2 zero_state[2] = {make_double2(1, 0), make_double2(0, 0)};
        for (int i = 0; i < n; ++i) {
            ensure_site_cap(i, bond_cap(i) * 2 * bond_cap(i + 1));
            HIPC(hipMemcpyAsync(sites[i].ptr, zero_state, 2 * sizeof(double2), hipMemcpyHostToDevice, stream));
            sites[i].l = sites[i].r = 1;
        }
        touched.assign(n, 0);
        sync();
        note_peak();
    }

    // ------------------------------------------------------------ one-qubit batch
    void run_one_qubit(const std::vector<PendingOne>& ops) {
        if (ops.empty()) return;
        tables_reset();
        std::vector<OneQubitJob> jobs(ops.size());
        int max_elems = 1;
        for (size_t i = 0; i < ops.size(); ++i) {
            wqc_tn_state* o = ops[i].owner ? ops[i].owner : this;
            Site& s = o->sites[ops[i].site];
            if (o != this) o->stats.one_qubit_gates++;
            jobs[i].site = s.ptr;
            jobs[i].l = s.l;
            jobs[i].r = s.r;
            for (int k = 0; k < 4; ++k) jobs[i].u[k] = ops[i].u[k];
            max_elems = std::max(max_elems, s.l * s.r);
        }
        if (tape)
            for (const auto& op : ops) {
                TapeOne t;
                t.site = op.site;
                t.tag = tape->tag;
                for (int k = 0; k < 4; ++k) t.u[k] = op.u[k];
                tape->order.emplace_back(0, (int)tape->ones.size());
                tape->ones.push_back(t);
            }
        OneQubitJob* d = upload(jobs);
        {
            Timer tm(this, 0);
            wqc::launch_one_qubit(d, (int)jobs.size(), max_elems, stream);
            check_launch();
            klaunch[0]++;
        }
        for (const auto& op : ops)
            if (!op.owner || op.owner == this) stats.one_qubit_gates++;
        sync();
        resolve_timings();
    }

    // ------------------------------------------------------------ timing helper
    // Brackets a group of launches on the state's stream with two events; the elapsed
    // time is read at the next synchronization (resolve_timings), so timing adds no
    // host waits of its own.
    hipEvent_t take_event() {
        if (!event_pool.empty()) {
            hipEvent_t e = event_pool.back();
            event_pool.pop_back();
            return e;
        }
        hipEvent_t e;
        HIPC(hipEventCreate(&e));
        return e;
    }
    struct Timer {
        wqc_tn_state* st;
        int fam;
        hipEvent_t a = nullptr;
        Timer(wqc_tn_state* s, int f) : st(s), fam(f) {
            if (!st->timing) return;
            a = st->take_event();
            HIPC(hipEventRecord(a, st->stream));
        }
        ~Timer() {
            if (!a) return;
            hipEvent_t b = st->take_event();
            if (hipEventRecord(b, st->stream) == hipSuccess) st->spans.push_back({a, b, fam});
        }
    };
    void resolve_timings() {
        for (auto& sp : spans) {
            float ms = 0;
            if (hipEventSynchronize(sp.b) == hipSuccess && hipEventElapsedTime(&ms, sp.a, sp.b) == hipSuccess)
                kms[sp.fam] += ms;
            event_pool.push_back(sp.a);
            event_pool.push_back(sp.b);
        }
        spans.clear();
    }

    // ------------------------------------------------------------ sweep graphs
    // The launches of one Jacobi sweep (flag clearing and every round of the batch) captured
    // once as a HIP graph and replayed for each sweep of the same form; the kernel arguments stay
    // fixed for the whole Jacobi call (the done flags are rewritten in place between sweeps).
    // WQC_TN_GRAPH=0 launches them one by one instead. The graph holds kernel nodes only.
    bool use_graph = true;
    struct SweepGraph {
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        SweepGraph() = default;
        SweepGraph(const SweepGraph&) = delete;
        SweepGraph& operator=(const SweepGraph&) = delete;
        ~SweepGraph() {
            if (exec) (void)hipGraphExecDestroy(exec);
            if (graph) (void)hipGraphDestroy(graph);
        }
    };
    template <class F>
    void run_sweep(SweepGraph& g, F&& launches) {
        if (!use_graph) {
            launches();
            check_launch();
            return;
        }
        if (!g.exec) {
            HIPC(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
            launches();
            const hipError_t le = hipGetLastError();
            hipGraph_t graph = nullptr;
            const hipError_t ce = hipStreamEndCapture(stream, &graph);
            g.graph = graph;
            HIPC(le);
            HIPC(ce);
            HIPC(hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0));
        }
        HIPC(hipGraphLaunch(g.exec, stream));
    }

    // ------------------------------------------------------------ Jacobi driver
    // Rotates the rows of every job's X until no block pair needs a rotation, then
    // writes the row norms into sigma (device, at sig_off[j]). Returns sweeps used.
    std::vector<int> last_sweeps, last_sweeps_lo;  // per job of the latest jacobi call

    // The jobs of a Jacobi call in groups that sweep independently, each on its own stream
    // (the state's stream and streams_x): a round waits for its slowest workgroup, and the
    // other groups' rounds fill the slots its early finishers leave. Each group has its own job
    // table, flags, pinned readback and sweep graphs; the host serves the groups in turn,
    // issuing a group's next sweep as soon as it has read that group's flags.
    struct SweepGroup {
        std::vector<int> ids;  // the caller's job indices
        hipStream_t s = nullptr;
        int nj = 0;
        const void* dj = nullptr;  // device job table of the group
        int* d_done = nullptr;
        int* d_rot = nullptr;
        unsigned long long* d_off = nullptr;
        char* h = nullptr;  // pinned readback: offmax (nj u64) then rotated (nj int)
        std::vector<int> done, jpairs;
        std::vector<double> prev_off;
        int active = 0, sweep = 0;
        bool in_flight = false;
        SweepGraph graphs[2];  // 256-thread and 512-thread rounds
        hipEvent_t ev = nullptr;
    };
    // Groups per Jacobi batch (WQC_TN_JACOBI_GROUPS, 1 to kMaxGroups). Measured on config 3:
    // FP64 form 1456 / 1389 / 1374 / 1357 ms per circuit with 1 / 2 / 3 / 4, mixed form 693 /
    // 690 / 690 / 689 ms.
    int jacobi_groups = kMaxGroups;
    // The mixed forms' low sweeps in one group (WQC_TN_JACOBI_GROUPS_LO, 1 to kMaxGroups): four
    // gave config 3 nothing (693 / 691 ms) and made the 256-slice batch slower (1.21 -> 1.32 s).
    int jacobi_groups_lo = 1;

    // Splits nj jobs (block pairs jp[j]) into min(jacobi_groups, nj) groups balanced by block
    // pairs (largest first, each to the lightest group). Sets up flags (in trace_buf,
    // cleared), readback regions and events.
    std::vector<std::unique_ptr<SweepGroup>> make_groups(const std::vector<int>& jp, size_t hread_off, int groups) {
        const int nj = (int)jp.size();
        const int ng = std::max(1, std::min(groups, nj));
        std::vector<std::unique_ptr<SweepGroup>> gs;
        for (int g = 0; g < ng; ++g) {
            gs.push_back(std::make_unique<SweepGroup>());
            gs.back()->s = g == 0 ? stream : streams_x[g - 1];
            gs.back()->ev = ev_sweep[g];
        }
        std::vector<int> order(nj);
        for (int j = 0; j < nj; ++j) order[j] = j;
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return jp[x] > jp[y]; });
        std::vector<long> load(ng, 0);
        for (int j : order) {
            int g = 0;
            for (int k = 1; k < ng; ++k)
                if (load[k] < load[g]) g = k;
            load[g] += jp[j];
            gs[g]->ids.push_back(j);
        }
        for (auto& g : gs) std::sort(g->ids.begin(), g->ids.end());
        // Flags: per group offmax (u64), rotated (int), done (int), 256-byte aligned.
        size_t off = 0;
        std::vector<size_t> fo(ng);
        for (int g = 0; g < ng; ++g) {
            fo[g] = off;
            const size_t n = gs[g]->ids.size();
            off += ((n * (sizeof(unsigned long long) + 2 * sizeof(int))) + 255) & ~(size_t)255;
        }
        trace_buf.ensure(off);
        note_peak();
        wqc::launch_fill_int(reinterpret_cast<int*>(trace_buf.p), 0, (int)(off / sizeof(int)), stream);
        size_t ho = hread_off;
        for (int g = 0; g < ng; ++g) {
            SweepGroup& G = *gs[g];
            G.nj = (int)G.ids.size();
            char* base = (char*)trace_buf.p + fo[g];
            G.d_off = reinterpret_cast<unsigned long long*>(base);
            G.d_rot = reinterpret_cast<int*>(G.d_off + G.nj);
            G.d_done = G.d_rot + G.nj;
            G.h = reinterpret_cast<char*>(hread) + ho;
            ho += ((size_t)G.nj * (sizeof(unsigned long long) + sizeof(int)) + 255) & ~(size_t)255;
            if (ho > hread_cap * sizeof(double)) fail(WQC_TN_EXEC_FAILED, "readback buffer too small");
            G.done.assign(G.nj, 0);
            G.prev_off.assign(G.nj, 0.0);
            G.active = G.nj;
            for (int k = 0; k < G.nj; ++k) G.jpairs.push_back(jp[G.ids[k]]);
        }
        return gs;
    }

    // Runs the groups' sweeps to the end: issue(g, wide) queues one sweep of group g on the
    // state's stream member (set to g's stream meanwhile) and its flag readback into g.h;
    // after(g) reads them and updates g.done / g.active. The other streams are forked from the
    // state's stream before and joined into it after.
    template <class Issue, class After>
    void sweep_groups(std::vector<std::unique_ptr<SweepGroup>>& gs, int max_sweeps, Issue&& issue, After&& after) {
        const hipStream_t main = stream;
        if (gs.size() > 1) {
            HIPC(hipEventRecord(ev_fork, main));
            for (auto& g : gs)
                if (g->s != main) HIPC(hipStreamWaitEvent(g->s, ev_fork, 0));
        }
        struct Restore {
            wqc_tn_state* st;
            hipStream_t s;
            ~Restore() { st->stream = s; }
        } restore{this, main};
        const int cus = dev->prop.multiProcessorCount;
        auto launch = [&](SweepGroup& g) {
            check_cancel();
            int wgs = 0;
            for (auto& o : gs)
                for (int k = 0; k < o->nj; ++k)
                    if (!o->done[k]) wgs += o->jpairs[k];
            const bool wide = use_wide && wgs <= cus;
            stream = g.s;
            issue(g, wide);
            HIPC(hipEventRecord(g.ev, g.s));
            stream = main;
            g.in_flight = true;
        };
        try {
            for (auto& g : gs)
                if (g->active > 0 && max_sweeps > 0) launch(*g);
            for (;;) {
                bool any = false;
                for (auto& gp : gs) {
                    SweepGroup& g = *gp;
                    if (!g.in_flight) continue;
                    any = true;
                    HIPC(hipEventSynchronize(g.ev));
                    g.in_flight = false;
                    stream = g.s;
                    after(g);
                    stream = main;
                    ++g.sweep;
                    if (g.active > 0 && g.sweep < max_sweeps) launch(g);
                }
                if (!any) break;
            }
        } catch (...) {
            // A cancellation or a failed call: let the other groups' queued sweeps end before
            // the state's buffers are used again.
            for (auto& g : gs)
                if (g->s != main) (void)hipStreamSynchronize(g->s);
            throw;
        }
        for (auto& g : gs)
            if (g->s != main) {
                HIPC(hipEventRecord(ev_join, g->s));
                HIPC(hipStreamWaitEvent(main, ev_join, 0));
            }
    }

    void jacobi(std::vector<JacobiJob>& jobs, double* d_sigma, const std::vector<long>& sig_off,
                int* d_flags) {
        (void)d_flags;
        const int nj = (int)jobs.size();
        last_sweeps.assign(nj, 0);
        if (!nj) return;
        JacobiJob* dj_all = upload(jobs);
        int max_n = 1;
        for (auto& j : jobs) max_n = std::max(max_n, j.n);
        // Per job: tolerance JACOBI_TOL sqrt(m), null-row floor NULL_ROWS_SQ / n (in the kernel).
        const double tol = JACOBI_TOL;
        const double nu2 = NULL_ROWS_SQ;
        unsigned long long* d_jstats = jstats_dev();
        // Largest relative off-diagonal |G_ij| / sqrt(G_ii G_jj) seen before rotation, per job
        // and sweep (always collected): a job whose value is at most stall_off and did not
        // halve since the previous sweep has reached the rounding floor of its rotations
        // (quadratic convergence would have taken it far below) and stops.
        // WQC_TN_JACOBI_TRACE=1 prints it per sweep for the widest jobs, and the history of
        // the jobs that needed the most sweeps.
        std::vector<std::vector<double>> hist(trace_jacobi ? nj : 0);
        // Block pairs per job: a sweep whose unfinished jobs have no more block pairs than the
        // device has CUs runs its rounds on the 512-thread form (one workgroup per CU either way,
        // and the round lasts as long as one workgroup).
        std::vector<int> jpairs(nj);
        for (int j = 0; j < nj; ++j) {
            const int b = (jobs[j].n + 15) / 16;
            jpairs[j] = (b + (b & 1)) / 2;
        }
        auto gs = make_groups(jpairs, 0, jacobi_groups);
        for (auto& g : gs) {
            std::vector<JacobiJob> sub;
            for (int j : g->ids) sub.push_back(jobs[j]);
            g->dj = upload(sub);
        }
        Timer tm(this, 3);
        sweep_groups(gs, MAX_SWEEPS, [&](SweepGroup& g, bool wide) {
            int gmax = 1;
            for (int j : g.ids) gmax = std::max(gmax, jobs[j].n);
            const int nb = (gmax + 15) / 16;
            const int nbe = nb + (nb & 1);
            const int rounds = nbe - 1, pairs = nbe / 2;
            const JacobiJob* dj = static_cast<const JacobiJob*>(g.dj);
            run_sweep(g.graphs[wide ? 1 : 0], [&] {
                wqc::launch_fill_int(reinterpret_cast<int*>(g.d_off), 0, 3 * g.nj, stream);
                for (int r = 0; r < rounds; ++r)
                    wqc::launch_jacobi_round(dj, g.nj, pairs, r, tol, nu2, inner_cap, g.d_rot, g.d_done, d_jstats,
                                             g.d_off, wide, stream);
            });
            klaunch[3] += rounds;
            stats.jacobi_sweeps += g.active;
            HIPC(hipMemcpyAsync(g.h, g.d_off, g.nj * (sizeof(unsigned long long) + sizeof(int)), hipMemcpyDeviceToHost,
                                stream));
        }, [&](SweepGroup& g) {
            const unsigned long long* offv = reinterpret_cast<const unsigned long long*>(g.h);
            const int* rot = reinterpret_cast<const int*>(offv + g.nj);
            if (trace_jacobi) {
                std::fprintf(stderr, "jacobi sweep %d stream %p active %d:", g.sweep + 1, (void*)g.s, g.active);
                int shown = 0;
                for (int k = 0; k < g.nj && shown < 4; ++k) {
                    const int j = g.ids[k];
                    if (g.done[k] || jobs[j].n < max_n) continue;
                    double v;
                    std::memcpy(&v, &offv[k], sizeof v);
                    std::fprintf(stderr, " [%dx%d] %.2e", jobs[j].n, jobs[j].m, v);
                    ++shown;
                }
                std::fprintf(stderr, "\n");
            }
            bool changed = false;
            for (int k = 0; k < g.nj; ++k) {
                if (g.done[k]) continue;
                const int j = g.ids[k];
                last_sweeps[j]++;
                double v;
                std::memcpy(&v, &offv[k], sizeof v);
                const bool stalled = rot[k] && v <= stall_off && g.prev_off[k] > 0.0 && v >= 0.5 * g.prev_off[k];
                g.prev_off[k] = v;
                if (trace_jacobi) hist[j].push_back(rot[k] ? v : 0.0);
                if (!rot[k] || stalled) {
                    g.done[k] = 1;
                    g.active--;
                    changed = true;
                    if (stalled) stats.jacobi_stalled++;
                }
            }
            // The rounds read the done flags at one address: the new flags are copied there.
            if (changed && g.active > 0)
                HIPC(hipMemcpyAsync(g.d_done, upload_ints(g.done), g.nj * sizeof(int), hipMemcpyDeviceToDevice, stream));
        });
        // A job still rotating after MAX_SWEEPS is counted; as the reference's Jacobi at its sweep
        // cap (src/tn/svd_guard.rs:165-229, then svd_reproduces), its factorization is kept when
        // it passes the norm and probe checks, which the caller applies to every split.
        for (auto& g : gs)
            for (int k = 0; k < g->nj; ++k)
                if (!g->done[k]) stats.jacobi_unconverged++;
        if (trace_jacobi) {
            std::vector<int> order(nj);
            for (int j = 0; j < nj; ++j) order[j] = j;
            std::sort(order.begin(), order.end(), [&](int a, int b) { return last_sweeps[a] > last_sweeps[b]; });
            for (int k = 0; k < std::min(nj, 3); ++k) {
                const int j = order[k];
                std::fprintf(stderr, "jacobi job [%dx%d] sweeps %d:", jobs[j].n, jobs[j].m, last_sweeps[j]);
                for (double v : hist[j]) std::fprintf(stderr, " %.1e", v);
                std::fprintf(stderr, "\n");
            }
        }
        std::vector<long> offs = sig_off;
        long* doff = upload(offs);
        wqc::launch_row_norms(dj_all, nj, max_n, d_sigma, doff, stream);
        check_launch();
    }

    // Mixed form of jacobi(): low-precision sweeps on complex64 rows with the rotation J
    // kept unitary in FP64 (k_jacobi_round_lo), then X <- J X_orig recomputed in FP64 and
    // polished by the FP64 rounds, which also produce the singular values.
    void jacobi_mixed(std::vector<JacobiJob>& jobs, double* d_sigma, const std::vector<long>& sig_off,
                      int* d_flags) {
        const int nj = (int)jobs.size();
        if (!nj) return;
        std::vector<int> lo_count(nj, 0);
        size_t off = 0;
        std::vector<size_t> o32(nj), o64(nj), oj32(nj), ow(nj), ojn(nj);
        auto take = [&](size_t bytes) {
            const size_t o = off;
            off = (off + bytes + 255) & ~(size_t)255;
            return o;
        };
        for (int k = 0; k < nj; ++k) {
            const size_t cnt = (size_t)jobs[k].n * jobs[k].m, nn = (size_t)jobs[k].n * jobs[k].n;
            o32[k] = take(cnt * sizeof(float2));
            o64[k] = take(cnt * sizeof(double2));
            oj32[k] = take(nn * sizeof(float2));
            ow[k] = take(nn * sizeof(double2));
            ojn[k] = take(nn * sizeof(double2));
        }
        ws32.ensure(off);
        note_peak();
        char* b32 = (char*)ws32.p;
        std::vector<const double2*> src(nj);
        std::vector<float2*> dst(nj);
        std::vector<long> counts(nj);
        std::vector<wqc::JacobiLoJob> lj(nj);
        long mx = 1;
        int max_n = 1;
        for (int k = 0; k < nj; ++k) {
            src[k] = jobs[k].X;
            dst[k] = reinterpret_cast<float2*>(b32 + o32[k]);
            counts[k] = (long)jobs[k].n * jobs[k].m;
            mx = std::max(mx, counts[k]);
            lj[k].X = dst[k];
            lj[k].J = reinterpret_cast<float2*>(b32 + oj32[k]);
            lj[k].n = jobs[k].n;
            lj[k].m = jobs[k].m;
            lj[k].ldx = jobs[k].m;
            lj[k].pad = 0;
            max_n = std::max(max_n, jobs[k].n);
        }
        wqc::launch_to_f32(upload(src), upload(dst), upload(counts), nj, mx, stream);
        // J32 = (complex64) J, the identity the split set up.
        {
            std::vector<const double2*> js(nj);
            std::vector<float2*> jd(nj);
            std::vector<long> jc(nj);
            long jm = 1;
            for (int k = 0; k < nj; ++k) {
                js[k] = jobs[k].J;
                jd[k] = lj[k].J;
                jc[k] = (long)jobs[k].n * jobs[k].n;
                jm = std::max(jm, jc[k]);
            }
            wqc::launch_to_f32(upload(js), upload(jd), upload(jc), nj, jm, stream);
        }
        check_launch();
        const float tol = (float)tol_lo;
        const double nu2 = NULL_ROWS_SQ;
        unsigned long long* d_jstats = jstats_dev();
        // Block pairs per job, for the groups and the choice of the 512-thread form (as in
        // jacobi()).
        std::vector<int> jpairs(nj);
        for (int j = 0; j < nj; ++j) {
            const int b = (jobs[j].n + 15) / 16;
            jpairs[j] = (b + (b & 1)) / 2;
        }
        auto gs = make_groups(jpairs, 0, jacobi_groups_lo);
        for (auto& g : gs) {
            std::vector<wqc::JacobiLoJob> sub;
            for (int j : g->ids) sub.push_back(lj[j]);
            g->dj = upload(sub);
        }
        const bool bf16 = precision == WQC_TN_PRECISION_MIXED_BF16;
        {
            Timer tm(this, 3);
            sweep_groups(gs, max_sweeps_lo, [&](SweepGroup& g, bool wide) {
                int gmax = 1;
                for (int j : g.ids) gmax = std::max(gmax, jobs[j].n);
                const int nb = (gmax + 15) / 16;
                const int nbe = nb + (nb & 1);
                const int rounds = nbe - 1, pairs = nbe / 2;
                const wqc::JacobiLoJob* dj = static_cast<const wqc::JacobiLoJob*>(g.dj);
                // offmax (collected only when tracing) and rotated, cleared together.
                unsigned long long* d_off = trace_jacobi ? g.d_off : nullptr;
                run_sweep(g.graphs[wide ? 1 : 0], [&] {
                    wqc::launch_fill_int(reinterpret_cast<int*>(g.d_off), 0, 3 * g.nj, stream);
                    for (int r = 0; r < rounds; ++r)
                        wqc::launch_jacobi_round_lo(dj, g.nj, pairs, r, tol, nu2, (float)floor_lo, inner_cap_lo,
                                                    g.d_rot, g.d_done, d_jstats, d_off, wide, bf16, stream);
                });
                klaunch[3] += rounds;
                stats.jacobi_sweeps_lo += g.active;
                HIPC(hipMemcpyAsync(g.h, g.d_off, g.nj * (sizeof(unsigned long long) + sizeof(int)),
                                    hipMemcpyDeviceToHost, stream));
            }, [&](SweepGroup& g) {
                const unsigned long long* offv = reinterpret_cast<const unsigned long long*>(g.h);
                const int* rot = reinterpret_cast<const int*>(offv + g.nj);
                if (trace_jacobi) {
                    std::fprintf(stderr, "low sweep %d stream %p active %d:", g.sweep + 1, (void*)g.s, g.active);
                    int shown = 0;
                    for (int k = 0; k < g.nj && shown < 4; ++k) {
                        const int j = g.ids[k];
                        if (g.done[k] || jobs[j].n < max_n) continue;
                        double v;
                        std::memcpy(&v, &offv[k], sizeof v);
                        std::fprintf(stderr, " [%dx%d] %.2e", jobs[j].n, jobs[j].m, v);
                        ++shown;
                    }
                    std::fprintf(stderr, "\n");
                }
                bool changed = false;
                for (int k = 0; k < g.nj; ++k) {
                    if (g.done[k]) continue;
                    lo_count[g.ids[k]]++;
                    if (!rot[k]) {
                        g.done[k] = 1;
                        g.active--;
                        changed = true;
                    }
                }
                if (changed && g.active > 0)
                    HIPC(hipMemcpyAsync(g.d_done, upload_ints(g.done), g.nj * sizeof(int), hipMemcpyDeviceToDevice,
                                        stream));
            });
        }
        // J = (complex128) J32, made unitary to FP64 precision by two Newton-Schulz steps
        // J <- J (3I - J^H J) / 2 (the FP32 rounds leave it unitary to about 1e-6).
        {
            Timer tm(this, 7);
            std::vector<const float2*> js(nj);
            std::vector<double2*> jd(nj);
            std::vector<long> jc(nj);
            long jm = 1;
            for (int k = 0; k < nj; ++k) {
                js[k] = lj[k].J;
                jd[k] = jobs[k].J;
                jc[k] = (long)jobs[k].n * jobs[k].n;
                jm = std::max(jm, jc[k]);
            }
            wqc::launch_to_f64(upload(js), upload(jd), upload(jc), nj, jm, stream);
            check_launch();
            for (int it = 0; it < 2; ++it) {
                std::vector<GemmJob> g1, g2;
                std::vector<const double*> cs;
                std::vector<double*> cd;
                std::vector<long> cc;
                for (int k = 0; k < nj; ++k) {
                    const int nn = jobs[k].n;
                    double2* src_j = it == 0 ? jobs[k].J : reinterpret_cast<double2*>(b32 + ojn[k]);
                    double2* dst_j = it == 0 ? reinterpret_cast<double2*>(b32 + ojn[k]) : jobs[k].J;
                    double2* w = reinterpret_cast<double2*>(b32 + ow[k]);
                    g1.push_back(gj(src_j, nn, 1, src_j, nn, 0, w, nn, nn, nn, nn, 1.0, 0.0));
                    cs.push_back(reinterpret_cast<const double*>(src_j));
                    cd.push_back(reinterpret_cast<double*>(dst_j));
                    cc.push_back(2L * nn * nn);
                    g2.push_back(gj(src_j, nn, 0, w, nn, 0, dst_j, nn, nn, nn, nn, -0.5, 1.5));
                }
                copy_batch(cs, cd, cc);
                gemm(g1, 7);
                gemm(g2, 7);
            }
        }
        // X <- J X in FP64 (through a scratch copy), then the FP64 rounds.
        {
            Timer tm(this, 7);
            std::vector<GemmJob> g;
            std::vector<const double*> cs;
            std::vector<double*> cd;
            std::vector<long> cc;
            for (int k = 0; k < nj; ++k) {
                double2* tmp = reinterpret_cast<double2*>(b32 + o64[k]);
                cs.push_back(reinterpret_cast<const double*>(jobs[k].X));
                cd.push_back(reinterpret_cast<double*>(tmp));
                cc.push_back(2 * counts[k]);
                g.push_back(gj(jobs[k].J, jobs[k].n, 0, tmp, jobs[k].m, 0, jobs[k].X, jobs[k].ldx, jobs[k].n,
                               jobs[k].m, jobs[k].n, 1, 0));
            }
            copy_batch(cs, cd, cc);
            gemm(g, 7);
        }
        jacobi(jobs, d_sigma, sig_off, d_flags);
        last_sweeps_lo = lo_count;
    }

    void jacobi_any(std::vector<JacobiJob>& jobs, double* d_sigma, const std::vector<long>& sig_off,
                    int* d_flags) {
        last_sweeps_lo.assign(jobs.size(), 0);
        if (precision != WQC_TN_PRECISION_FP64)
            jacobi_mixed(jobs, d_sigma, sig_off, d_flags);
        else
            jacobi(jobs, d_sigma, sig_off, d_flags);
    }

    int* upload_ints(const std::vector<int>& v) { return upload(v); }

    // Every kept job's row permutation in one upload; fj[q].perm points at perms[q].
    void set_perms(std::vector<FactorJob>& fj, const std::vector<std::vector<int>>& perms) {
        std::vector<int> all;
        std::vector<size_t> at(fj.size());
        for (size_t q = 0; q < fj.size(); ++q) {
            at[q] = all.size();
            all.insert(all.end(), perms[q].begin(), perms[q].end());
        }
        int* d = upload(all);
        for (size_t q = 0; q < fj.size(); ++q) fj[q].perm = d + at[q];
    }

    // Device-to-device copies of a batch (counts in doubles) as one launch.
    void copy_batch(const std::vector<const double*>& src, const std::vector<double*>& dst,
                    const std::vector<long>& counts) {
        if (src.empty()) return;
        long mx = 1;
        std::vector<wqc::CopyJob> jobs(src.size());
        for (size_t i = 0; i < src.size(); ++i) {
            jobs[i] = wqc::CopyJob{src[i], dst[i], counts[i], 0};
            mx = std::max(mx, counts[i]);
        }
        wqc::launch_copy_batch(upload(jobs), (int)jobs.size(), mx, stream);
        check_launch();
    }

    // Device counters of the Jacobi rounds: [0] inner sweeps, [1] pair eigensolves.
    DevBuf jstats_buf;
    unsigned long long* jstats_dev() {
        if (!jstats_buf.p) {
            jstats_buf.ensure(64);
            HIPC(hipMemsetAsync(jstats_buf.p, 0, 64, stream));
        }
        return reinterpret_cast<unsigned long long*>(jstats_buf.p);
    }
    void read_jstats() {
        if (!jstats_buf.p) return;
        unsigned long long v[2] = {0, 0};
        HIPC(hipMemcpyAsync(v, jstats_buf.p, sizeof v, hipMemcpyDeviceToHost, stream));
        sync();
        stats.jacobi_inner_sweeps = v[0];
        stats.jacobi_pair_solves = v[1];
    }

    // Frobenius^2 (deterministic two-stage sum) then checked_svd's dust rule, per matrix.
    void dust(const std::vector<double2*>& mats, const std::vector<long>& counts, double* d_fro2,
              double* d_partial, int nblk) {
        const int nj = (int)mats.size();
        if (!nj) return;
        double2** dp = upload(mats);
        long* dc = upload(counts);
        long mx = 1;
        for (long c : counts) mx = std::max(mx, c);
        wqc::launch_sumsq(dp, dc, d_partial, nj, nblk, stream);
        wqc::launch_sumsq_final(d_partial, d_fro2, nj, nblk, stream);
        wqc::launch_dust(dp, dc, d_fro2, DUST_REL, nj, mx, stream);
        check_launch();
    }

    // ------------------------------------------------------------ two-site batch
    struct JobInfo {
        wqc_tn_state* st;  // the state whose sites the unit updates
        int p, dl, bond, dr, rows, cols;
        bool rsvd;
        size_t ell;
        // workspace offsets (bytes)
        size_t o_theta, o_thetaT, o_J, o_sig;
        size_t o_Y0, o_Z, o_Y, o_V, o_T, o_Z2, o_B, o_JB, o_sigB, o_W, o_W2;
        size_t o_lqU = 0, o_lqUT = 0, o_lqUTh = 0, o_lqR = 0, o_lqW = 0, o_lqT = 0, o_lqG = 0;  // LQ preconditioning
        size_t o_lq2U = 0, o_lq2UT = 0, o_lq2UTh = 0, o_lq2R = 0;  // its second factorization (of R)
        OmegaEntry* om;
        size_t chi_kept;
        bool fallback;
        int sweeps = 0, sweeps_lo = 0;
        double dw = 0.0;  // discarded weight sum_{k >= chi} s_k^2 / sum s_k^2 of the split
        double probe = 0.0;  // max_p |A x_p - U S V^H x_p| / (|A|_F |x_p|) of the kept factorization
        int rec = -1;        // its entry in the replay tape
    };

    void run_units(std::vector<Unit>& units) {
        if (units.empty()) return;
        check_cancel();
        sync();
        tables_reset();
        const int nu = (int)units.size();
        std::vector<JobInfo> J(nu);
        size_t off = 0;
        auto take = [&](size_t bytes) {
            size_t o = (off + 255) & ~(size_t)255;
            off = o + bytes;
            return o;
        };
        const size_t C16 = sizeof(double2);
        for (int i = 0; i < nu; ++i) {
            JobInfo& j = J[i];
            j.st = units[i].owner ? units[i].owner : this;
            j.p = units[i].p;
            const Site& L = j.st->sites[j.p];
            const Site& R = j.st->sites[j.p + 1];
            if (L.r != R.l) fail(WQC_TN_EXEC_FAILED, "MPS bond layout mismatch between sites");
            j.dl = L.l;
            j.bond = L.r;
            j.dr = R.r;
            j.rows = 2 * j.dl;
            j.cols = 2 * j.dr;
            const size_t rows = j.rows, cols = j.cols;
            j.ell = std::min(std::min(j.st->chi + RSVD_OVERSAMPLE, rows), cols);
            j.rsvd = !j.st->reject && !j.st->exact_only &&
                     !(j.ell == 0 || j.ell >= std::min(rows, cols) || !rsvd_worth(rows, cols, j.ell));
            j.om = nullptr;
            j.fallback = false;
            const size_t n_ex = std::min(rows, cols);
            j.o_theta = take(rows * cols * C16);
            j.o_J = take(n_ex * n_ex * C16);
            j.o_sig = take(std::max(rows, cols) * sizeof(double));
            if (use_lq && n_ex >= (size_t)LQ_MIN_N && std::max(rows, cols) <= (size_t)wqc::LQ_MAX_COLS) {
                j.o_lqU = take(rows * cols * C16);
                j.o_lqUT = take(rows * cols * C16);
                j.o_lqUTh = take(rows * cols * C16);
                j.o_lqR = take(n_ex * n_ex * C16);
                j.o_lqW = take(n_ex * wqc::LQ_SB * C16);
                j.o_lqT = take((size_t)wqc::LQ_SB * wqc::LQ_SB * C16);
                j.o_lqG = take((size_t)wqc::LQ_SB * wqc::LQ_SB * C16);
                if (use_lq2) {
                    j.o_lq2U = take(n_ex * n_ex * C16);
                    j.o_lq2UT = take(n_ex * n_ex * C16);
                    j.o_lq2UTh = take(n_ex * n_ex * C16);
                    j.o_lq2R = take(n_ex * n_ex * C16);
                }
            }
            if (j.rsvd) {
                j.om = &omega_for(dev, rows, cols, j.ell);
                const size_t r = std::max(1, j.om->r);
                j.o_thetaT = rows > cols ? take(rows * cols * C16) : 0;
                j.o_Y0 = take(rows * r * C16);
                j.o_Z = take(cols * r * C16);
                j.o_Y = take(rows * r * C16);
                j.o_V = take(rows * r * C16);
                j.o_T = take(r * r * C16);
                j.o_Z2 = take(r * cols * C16);
                j.o_B = take(j.ell * cols * C16);
                j.o_JB = take(j.ell * j.ell * C16);
                j.o_sigB = take(j.ell * sizeof(double));
                j.o_W = take(j.ell * std::min(j.st->chi, j.ell) * C16);
                j.o_W2 = take(r * std::min(j.st->chi, j.ell) * C16 * 2);
            }
        }
        const int nblk = 64;
        const size_t o_flags = take(4 * (size_t)nu * sizeof(int));
        const size_t o_fro = take((size_t)nu * 2 * sizeof(double));
        const size_t o_part = take((size_t)nu * 2 * nblk * sizeof(double));
        ws.ensure(off);
        note_peak();
        char* base = (char*)ws.p;
        auto P = [&](size_t o) { return reinterpret_cast<double2*>(base + o); };
        auto PD = [&](size_t o) { return reinterpret_cast<double*>(base + o); };
        int* d_flags = reinterpret_cast<int*>(base + o_flags);
        double* d_fro = PD(o_fro);
        double* d_part = PD(o_part);

        if (tape) {
            if (use_lq || precision != WQC_TN_PRECISION_FP64)
                fail(WQC_TN_EXEC_FAILED, "the adjoint replay records plain FP64 splits only");
            for (int i = 0; i < nu; ++i) {
                JobInfo& j = J[i];
                if (j.rsvd || j.st != this) fail(WQC_TN_EXEC_FAILED, "the adjoint replay records exact splits only");
                TapeUnit t;
                t.p = j.p;
                t.dl = j.dl;
                t.bond = j.bond;
                t.dr = j.dr;
                t.tag = tape->tag;
                std::memcpy(t.u4, units[i].u4, sizeof(t.u4));
                const size_t nl = (size_t)j.dl * 2 * j.bond, nr = (size_t)j.bond * 2 * j.dr;
                t.L = tape->arena.take(nl);
                t.R = tape->arena.take(nr);
                HIPC(hipMemcpyAsync(t.L, sites[j.p].ptr, nl * sizeof(double2), hipMemcpyDeviceToDevice, stream));
                HIPC(hipMemcpyAsync(t.R, sites[j.p + 1].ptr, nr * sizeof(double2), hipMemcpyDeviceToDevice, stream));
                j.rec = (int)tape->units.size();
                tape->order.emplace_back(1, j.rec);
                tape->units.push_back(std::move(t));
            }
        }
        // 1. merge + U (exact jobs with rows > cols store theta'^H for the row Jacobi).
        {
            Timer tm(this, 1);
            std::vector<MergeJob> mj(nu);
            int mdl = 1, mdr = 1;
            for (int i = 0; i < nu; ++i) {
                JobInfo& j = J[i];
                mj[i].L = j.st->sites[j.p].ptr;
                mj[i].R = j.st->sites[j.p + 1].ptr;
                mj[i].out = P(j.o_theta);
                mj[i].dl = j.dl;
                mj[i].bond = j.bond;
                mj[i].dr = j.dr;
                mj[i].transpose = (!j.rsvd && j.rows > j.cols) ? 1 : 0;
                std::memcpy(mj[i].u4, units[i].u4, sizeof(mj[i].u4));
                mdl = std::max(mdl, j.dl);
                mdr = std::max(mdr, j.dr);
            }
            wqc::launch_merge(upload(mj), nu, mdl, mdr, stream);
            check_launch();
            klaunch[1]++;
        }
        for (const JobInfo& j : J) j.st->stats.two_qubit_units++;

        SplitBatch rs, ex;
        for (int i = 0; i < nu; ++i) (J[i].rsvd ? rs.ids : ex.ids).push_back(i);
        rs.d_fro = d_fro;
        rs.d_part = d_part;
        ex.d_fro = d_fro + nu;
        ex.d_part = d_part + (size_t)nu * nblk;
        if (!rs.ids.empty()) prep_rsvd(J, rs, base, nblk);
        if (!ex.ids.empty()) prep_exact(J, ex, base, nblk);
        jacobi_splits(J, {&rs, &ex}, base, d_flags);
        if (!rs.ids.empty()) finish_rsvd(J, rs, base);
        if (!ex.ids.empty()) finish_exact(J, ex, base);
        SplitBatch fb;
        for (int i : rs.ids)
            if (J[i].fallback) fb.ids.push_back(i);
        if (!fb.ids.empty()) {
            fb.d_fro = d_fro + nu;
            fb.d_part = d_part + (size_t)nu * nblk;
            prep_exact(J, fb, base, nblk);
            jacobi_splits(J, {&fb}, base, d_flags);
            finish_exact(J, fb, base);
        }
        sync();
        resolve_timings();
        for (int i = 0; i < nu; ++i)
            if (J[i].st->split_log_on) {
                const JobInfo& j = J[i];
                const int path = !j.rsvd ? 0 : (j.fallback ? 2 : 1);
                const int32_t e[8] = {j.rows, j.cols, j.bond, path, (int32_t)j.ell, (int32_t)j.chi_kept,
                                      j.sweeps, j.sweeps_lo};
                j.st->split_log.insert(j.st->split_log.end(), e, e + 8);
                j.st->split_dw.push_back(j.dw);
                j.st->split_probe.push_back(j.probe);
            }
    }

    // Batched GEMM. Products with few output tiles and a long K (the LQ's panel and block
    // products) are split over K into up to 8 slices, summed in slice order by a second kernel.
    DevBuf splitk_buf;
    // One batched product: its jobs sit at off in a job table, with the grid bounds and the
    // split-K factor. Every split-K call reuses the scratch from its start (the calls run in
    // stream order, each reduced before the next begins).
    struct GemmCall {
        size_t off = 0;
        int n = 0, mm = 1, mn = 1, splits = 1;
        int tile = 64;               // the autotuner's choice for this shape class (or 64)
        bool tuned = false;          // the class has a measured tile
        std::array<int, 6> key{};    // shape class
    };
    static int log2_bucket(int v) {
        int b = 0;
        while ((1 << b) < v && b < 30) ++b;
        return b;
    }
    GemmCall size_gemm(const std::vector<GemmJob>& g, size_t& need) const {
        GemmCall c;
        int mk = 1;
        for (auto& x : g) {
            c.mm = std::max(c.mm, x.M);
            c.mn = std::max(c.mn, x.N);
            mk = std::max(mk, x.K);
        }
        c.n = (int)g.size();
        const long tiles = (long)((c.mm + 63) / 64) * ((c.mn + 63) / 64);
        c.splits = (use_splitk && tiles <= 16 && mk >= 256) ? std::min(8, mk / 128) : 1;
        if (c.splits > 1) {
            size_t nd = 0;
            for (auto& x : g) nd += (size_t)c.splits * x.M * x.N;
            need = std::max(need, nd);
        }
        c.key = {log2_bucket(c.mm), log2_bucket(c.mn), log2_bucket(mk), log2_bucket(c.n), c.splits,
                 g.empty() ? 0 : g[0].opA * 2 + g[0].opB};
        std::lock_guard<std::mutex> lk(dev->mu);
        auto it = dev->gemm_tile.find(c.key);
        if (it != dev->gemm_tile.end()) {
            c.tile = it->second;
            c.tuned = true;
        }
        return c;
    }
    // A product may be timed (run more than once) only when it overwrites its output: every job
    // with beta = 0 and an output that is none of its inputs.
    static bool rerunnable(const std::vector<GemmJob>& g) {
        for (const auto& x : g)
            if (x.beta.x != 0.0 || x.beta.y != 0.0 || x.C == x.A || x.C == x.B) return false;
        return true;
    }
    bool autotune = true;  // WQC_TN_AUTOTUNE=0: no measurements (cached or default tiles)
    int forced_tile = 0;   // WQC_TN_GEMM_TILE=32|64: that tile for every product, no tuning
    // The autotuner: both tiles run twice on the call's own (rerunnable) inputs, alternating, and
    // the faster minimum is kept for the shape class. Both tiles give the same bits, so the
    // outputs are valid whichever ran last and the choice never changes a result.
    void tune_call(const GemmJob* table, GemmCall& c, int family) {
        hipEvent_t ev[5];
        for (auto& e : ev) e = take_event();
        HIPC(hipEventRecord(ev[0], stream));
        for (int rep = 0; rep < 2; ++rep)
            for (int t = 0; t < 2; ++t) {
                GemmCall x = c;
                x.tile = t == 0 ? 64 : 32;
                launch_one(table, x, family);
                HIPC(hipEventRecord(ev[1 + 2 * rep + t], stream));
            }
        HIPC(hipEventSynchronize(ev[4]));
        float ms[4] = {0, 0, 0, 0};
        for (int k = 0; k < 4; ++k) HIPC(hipEventElapsedTime(&ms[k], ev[k], ev[k + 1]));
        const float t64 = std::min(ms[0], ms[2]), t32 = std::min(ms[1], ms[3]);
        for (auto& e : ev) event_pool.push_back(e);
        c.tile = t32 < t64 ? 32 : 64;
        c.tuned = true;
        std::lock_guard<std::mutex> lk(dev->mu);
        if (dev->gemm_tile.emplace(c.key, c.tile).second) dev->save_tune_entry(c.key, c.tile);
    }
    void point_parts(std::vector<GemmJob>& g, const GemmCall& c) {
        if (c.splits <= 1) return;
        double2* p = reinterpret_cast<double2*>(splitk_buf.p);
        for (auto& x : g) {
            x.part = p;
            p += (size_t)c.splits * x.M * x.N;
        }
    }
    void launch_one(const GemmJob* table, const GemmCall& c, int family) {
        if (c.splits > 1)
            wqc::launch_gemm_splitk(table + c.off, c.n, c.mm, c.mn, c.splits, stream, c.tile);
        else
            wqc::launch_gemm(table + c.off, c.n, c.mm, c.mn, stream, c.tile);
        check_launch();
        klaunch[family]++;
    }
    // Launches a call with its class's tile; an untuned class whose jobs may be rerun is measured
    // first (its timing runs leave the product in place).
    void launch_call(const GemmJob* table, GemmCall c, int family, const std::vector<GemmJob>* host = nullptr) {
        if (c.n == 0) return;
        if (forced_tile) {
            c.tile = forced_tile;
            launch_one(table, c, family);
            return;
        }
        if (!c.tuned && autotune && host && rerunnable(*host)) {
            tune_call(table, c, family);
            return;
        }
        launch_one(table, c, family);
    }
    void gemm(std::vector<GemmJob>& g, int family = 2) {
        if (g.empty()) return;
        size_t need = 0;
        const GemmCall c = size_gemm(g, need);
        if (need) {
            splitk_buf.ensure(need * sizeof(double2));
            note_peak();
        }
        point_parts(g, c);
        launch_call(upload(g), c, family, &g);
    }
    // A fixed sequence of products (and other launches between them, run by `between`):
    // every call's jobs go up in one table upload instead of one per call.
    void gemm_sequence(std::vector<std::vector<GemmJob>>& calls, int family,
                       const std::function<void(size_t)>& between) {
        size_t need = 0;
        std::vector<GemmCall> cs(calls.size());
        for (size_t i = 0; i < calls.size(); ++i) cs[i] = size_gemm(calls[i], need);
        if (need) {
            splitk_buf.ensure(need * sizeof(double2));
            note_peak();
        }
        std::vector<GemmJob> all;
        for (size_t i = 0; i < calls.size(); ++i) {
            point_parts(calls[i], cs[i]);
            cs[i].off = all.size();
            all.insert(all.end(), calls[i].begin(), calls[i].end());
        }
        const GemmJob* table = all.empty() ? nullptr : upload(all);
        for (size_t i = 0; i < calls.size(); ++i) {
            between(i);
            launch_call(table, cs[i], family, &calls[i]);
        }
        between(calls.size());
    }

    static GemmJob gj(const double2* A, int lda, int opA, const double2* B, int ldb, int opB, double2* C,
                      int ldc, int M, int N, int K, double alpha, double beta) {
        GemmJob g;
        g.A = A;
        g.B = B;
        g.C = C;
        g.M = M;
        g.N = N;
        g.K = K;
        g.lda = lda;
        g.ldb = ldb;
        g.ldc = ldc;
        g.opA = opA;
        g.opB = opB;
        g.alpha = make_double2(alpha, 0);
        g.beta = make_double2(beta, 0);
        g.part = nullptr;
        return g;
    }

    // A batch of splits of one kind: its jobs, their Jacobi problems (rows of theta' or of the
    // range finder's B), sigma offsets, |A|_F^2 slots and, after jacobi_splits, where the
    // host copies of sigma and |A|_F^2 sit in the readback buffer.
    struct SplitBatch {
        std::vector<int> ids;
        std::vector<JacobiJob> jj;
        // LQ-preconditioned jobs: jj[k].X is R; orig[k] (row stride orig_ld[k]) held X and,
        // after the Jacobi, holds M = [J 0] Q^H.
        std::vector<char> pre;
        std::vector<double2*> orig;
        std::vector<int> orig_ld;
        std::vector<long> offs;
        double* d_fro = nullptr;
        double* d_part = nullptr;
        std::vector<size_t> hoff;
        size_t hfro = 0;
        // Column count of each job's matrix before any preconditioning, and where the two
        // squared probe residuals of job k sit in the readback (hprobe + 2k, +1).
        std::vector<int> m0;
        size_t hprobe = 0;
    };

    // The probe half of svd_reproduces (src/tn/svd_guard.rs:261-281): two fixed vectors
    // x_p[j] = (sin t, cos t), t = (j + 1)(0.754877666 + 0.56984029 p), |x_p[j]| = 1, as the
    // columns of P (m x 2 row-major). A factorization passes when |A x_p - U S V^H x_p| <=
    // 1e-9 |A|_F |x_p| + 1e-14 for both, A being the matrix the Jacobi takes (after the dust
    // rule: theta', its adjoint when tall, or the range finder's B).
    DevBuf probe_buf, probe_p_buf;
    int probe_m = 0;
    double2* probe_matrix(int m) {
        if (m > probe_m) {
            const int cap = std::max(m, 1024);
            std::vector<double2> P((size_t)cap * 2);
            for (int j = 0; j < cap; ++j)
                for (int p = 0; p < 2; ++p) {
                    const double t = (j + 1.0) * (0.754877666 + p * 0.56984029);
                    P[(size_t)j * 2 + p] = make_double2(std::sin(t), std::cos(t));
                }
            sync();
            probe_p_buf.ensure(P.size() * sizeof(double2));
            HIPC(hipMemcpyAsync(probe_p_buf.p, P.data(), P.size() * sizeof(double2), hipMemcpyHostToDevice, stream));
            sync();
            probe_m = cap;
            note_peak();
        }
        return reinterpret_cast<double2*>(probe_p_buf.p);
    }

    // The larger relative probe residual of job k of a finished batch (|A|_F^2 = fro2).
    double probe_rel(const SplitBatch& b, int k, double fro2) const {
        const double scale = std::sqrt(std::max(fro2, 2.2250738585072014e-308));
        const double xn = std::sqrt((double)b.m0[k]);
        const double e = std::max(hread[b.hprobe + 2 * (size_t)k], hread[b.hprobe + 2 * (size_t)k + 1]);
        return std::sqrt(e) / (scale * xn);
    }

    // Whether job k of a finished batch passes the probe check (|A|_F^2 = fro2).
    bool probes_pass(const SplitBatch& b, int k, double fro2) const {
        const double scale = std::sqrt(std::max(fro2, 2.2250738585072014e-308));
        const double xn = std::sqrt((double)b.m0[k]);
        for (int p = 0; p < 2; ++p) {
            const double err = std::sqrt(hread[b.hprobe + 2 * (size_t)k + p]);
            if (!(err <= SVD_CHECK_TOL * scale * xn + 1e-14)) return false;
        }
        return true;
    }

    // Exact split (truncated_svd_split): theta' (or its adjoint when tall) as the Jacobi rows,
    // the dust rule, J = I.
    void prep_exact(std::vector<JobInfo>& J, SplitBatch& b, char* base, int nblk) {
        auto P = [&](size_t o) { return reinterpret_cast<double2*>(base + o); };
        auto PD = [&](size_t o) { return reinterpret_cast<double*>(base + o); };
        const std::vector<int>& ids = b.ids;
        const int ne = (int)ids.size();
        // Fallback jobs from the range finder hold theta' untransposed.
        for (int i : ids) {
            JobInfo& j = J[i];
            if (j.rsvd && j.rows > j.cols) {
                wqc::launch_conj_transpose(P(j.o_theta), P(j.o_thetaT), j.rows, j.cols, stream);
                check_launch();
            }
        }
        std::vector<JacobiJob>& jj = b.jj;
        jj.assign(ne, JacobiJob{});
        std::vector<double2*> mats(ne), Jm(ne);
        std::vector<long> counts(ne);
        std::vector<long>& offs = b.offs;
        offs.assign(ne, 0);
        std::vector<int> dims(ne);
        int max_n = 1;
        for (int k = 0; k < ne; ++k) {
            JobInfo& j = J[ids[k]];
            const bool tall = j.rows > j.cols;
            double2* X = (j.rsvd && tall) ? P(j.o_thetaT) : P(j.o_theta);
            jj[k].X = X;
            jj[k].J = P(j.o_J);
            jj[k].n = tall ? j.cols : j.rows;
            jj[k].m = tall ? j.rows : j.cols;
            jj[k].ldx = jj[k].m;
            mats[k] = X;
            Jm[k] = P(j.o_J);
            counts[k] = (long)j.rows * j.cols;
            dims[k] = jj[k].n;
            max_n = std::max(max_n, jj[k].n);
            offs[k] = (long)((j.o_sig) / sizeof(double));
        }
        dust(mats, counts, b.d_fro, b.d_part, nblk);
        wqc::launch_identity(upload(Jm), upload(dims), ne, max_n, stream);
        check_launch();
    }

    // Range finder (rsvd_sites): Y = A (A^H (A Omega_I)), compact-WY Householder of its
    // independent columns, B = Q^H A; B's rows are the Jacobi rows.
    void prep_rsvd(std::vector<JobInfo>& J, SplitBatch& b, char* base, int nblk) {
        auto P = [&](size_t o) { return reinterpret_cast<double2*>(base + o); };
        auto PD = [&](size_t o) { return reinterpret_cast<double*>(base + o); };
        const std::vector<int>& ids = b.ids;
        const int nr = (int)ids.size();
        {
            Timer tm(this, 2);
            std::vector<GemmJob> g1, g2, g3;
            for (int i : ids) {
                JobInfo& j = J[i];
                const int r = j.om->r;
                // Y0 = A Omega_I; Z = A^H Y0; Y = A Z
                g1.push_back(gj(P(j.o_theta), j.cols, 0, j.om->omega_cols, r, 0, P(j.o_Y0), r, j.rows, r, j.cols, 1, 0));
                g2.push_back(gj(P(j.o_theta), j.cols, 1, P(j.o_Y0), r, 0, P(j.o_Z), r, j.cols, r, j.rows, 1, 0));
                g3.push_back(gj(P(j.o_theta), j.cols, 0, P(j.o_Z), r, 0, P(j.o_Y), r, j.rows, r, j.cols, 1, 0));
            }
            gemm(g1);
            gemm(g2);
            gemm(g3);
            std::vector<HouseJob> hj;
            for (int i : ids) {
                JobInfo& j = J[i];
                if (j.om->r > wqc::kMaxHouseCols) fail(WQC_TN_EXEC_FAILED, "range finder rank above the QR limit");
                HouseJob h;
                h.Y = P(j.o_Y);
                h.V = P(j.o_V);
                h.T = P(j.o_T);
                h.pos = j.om->pos;
                h.rows = j.rows;
                h.r = j.om->r;
                h.pad0 = h.pad1 = 0;
                h.tau = QR_RANK_TOL;
                hj.push_back(h);
            }
            wqc::launch_householder(upload(hj), (int)hj.size(), stream);
            check_launch();
            klaunch[2]++;
            // B = A[0:ell, :] - V[0:ell, :] (T^H (V^H A))
            std::vector<GemmJob> g4, g5, g6;
            for (int i : ids) {
                JobInfo& j = J[i];
                const int r = j.om->r;
                g4.push_back(gj(P(j.o_V), r, 1, P(j.o_theta), j.cols, 0, P(j.o_Z2), j.cols, r, j.cols, j.rows, 1, 0));
                g5.push_back(gj(P(j.o_T), r, 1, P(j.o_Z2), j.cols, 0, P(j.o_Z), j.cols, r, j.cols, r, 1, 0));
                HIPC(hipMemcpyAsync(P(j.o_B), P(j.o_theta), j.ell * (size_t)j.cols * sizeof(double2),
                                    hipMemcpyDeviceToDevice, stream));
                g6.push_back(gj(P(j.o_V), r, 0, P(j.o_Z), j.cols, 0, P(j.o_B), j.cols, (int)j.ell, j.cols, r, -1, 1));
            }
            gemm(g4);
            gemm(g5);
            gemm(g6);
        }
        // checked_svd(B): dust rule, Jacobi on the rows of B (ell <= cols), norms.
        std::vector<JacobiJob>& jj = b.jj;
        jj.assign(nr, JacobiJob{});
        std::vector<double2*> mats(nr), Jm(nr);
        std::vector<long> counts(nr);
        std::vector<long>& offs = b.offs;
        offs.assign(nr, 0);
        std::vector<int> dims(nr);
        int max_n = 1;
        for (int k = 0; k < nr; ++k) {
            JobInfo& j = J[ids[k]];
            jj[k].X = P(j.o_B);
            jj[k].J = P(j.o_JB);
            jj[k].n = (int)j.ell;
            jj[k].m = j.cols;
            jj[k].ldx = j.cols;
            mats[k] = P(j.o_B);
            Jm[k] = P(j.o_JB);
            counts[k] = (long)j.ell * j.cols;
            dims[k] = (int)j.ell;
            max_n = std::max(max_n, (int)j.ell);
            offs[k] = (long)(j.o_sigB / sizeof(double));
        }
        dust(mats, counts, b.d_fro, b.d_part, nblk);
        wqc::launch_identity(upload(Jm), upload(dims), nr, max_n, stream);
        check_launch();
    }

    // Householder LQ of every eligible Jacobi job (X = L Q, panels of LQ_NB rows, trailing
    // rows updated by two GEMMs per panel); the job then rotates the rows of R = L^H.
    void lq_precondition(std::vector<JobInfo>& J, const std::vector<SplitBatch*>& batches, char* base,
                         std::vector<wqc::LqJob>& lq, std::vector<double2*>& lqw, std::vector<wqc::LqJob>& lq2,
                         std::vector<double2*>& lqw2) {
        auto P = [&](size_t o) { return reinterpret_cast<double2*>(base + o); };
        lq.clear();
        lqw.clear();
        lq2.clear();
        lqw2.clear();
        for (auto* b : batches) {
            const size_t nb = b->ids.size();
            b->pre.assign(nb, 0);
            b->orig.assign(nb, nullptr);
            b->orig_ld.assign(nb, 0);
            for (size_t i = 0; i < nb; ++i) {
                JobInfo& j = J[b->ids[i]];
                JacobiJob& x = b->jj[i];
                if (!j.o_lqU || x.n < LQ_MIN_N || x.m > wqc::LQ_MAX_COLS || x.n > x.m) continue;
                wqc::LqJob q;
                q.X = x.X;
                q.Ubc = P(j.o_lqU);
                q.UTc = P(j.o_lqUT);
                q.UThc = P(j.o_lqUTh);
                q.R = P(j.o_lqR);
                q.J = x.J;
                q.T64 = P(j.o_lqT);
                q.Gu = P(j.o_lqG);
                q.n = x.n;
                q.m = x.m;
                q.ldx = x.ldx;
                q.pad = 0;
                lq.push_back(q);
                lqw.push_back(P(j.o_lqW));
                b->pre[i] = 1;
                b->orig[i] = x.X;
                b->orig_ld[i] = x.ldx;
                x.X = q.R;
                x.m = x.n;
                x.ldx = x.n;
            }
        }
        if (lq.empty()) return;
        lq_run(lq, lqw);
        // Second factorization R = L2 Q2 of every preconditioned job, the Jacobi then rotating
        // the rows of R2 = L2^H (n x n) instead of R's: R2 is nearer diagonal (a further step of
        // the unshifted QR iteration), which on the splits with a graded tail of singular values
        // through the cutoff takes 6 FP64 sweeps instead of 18. Every preconditioned job has
        // its buffers, so lq2[k] belongs to lq[k].
        if (!use_lq2) return;
        for (auto* b : batches)
            for (size_t i = 0; i < b->ids.size(); ++i) {
                if (!b->pre[i]) continue;
                JobInfo& j = J[b->ids[i]];
                JacobiJob& x = b->jj[i];
                if (!j.o_lq2U) fail(WQC_TN_EXEC_FAILED, "second LQ factorization without its buffers");
                wqc::LqJob q;
                q.X = x.X;  // R, factored in place
                q.Ubc = P(j.o_lq2U);
                q.UTc = P(j.o_lq2UT);
                q.UThc = P(j.o_lq2UTh);
                q.R = P(j.o_lq2R);
                q.J = x.J;
                q.T64 = P(j.o_lqT);
                q.Gu = P(j.o_lqG);
                q.n = x.n;
                q.m = x.n;
                q.ldx = x.ldx;
                q.pad = 0;
                lq2.push_back(q);
                lqw2.push_back(P(j.o_lqW));
                x.X = q.R;
            }
        lq_run(lq2, lqw2);
    }

    // The factorization's fixed sequence of launches for the jobs of lq (each X = L Q in place,
    // R = L^H extracted).
    void lq_run(const std::vector<wqc::LqJob>& lq, const std::vector<double2*>& lqw) {
        Timer tm(this, 6);
        int max_n = 0;
        for (const auto& q : lq) max_n = std::max(max_n, q.n);
        const int nl = (int)lq.size();
        wqc::LqJob* dl = upload(lq);
        const int NB = wqc::LQ_NB, SB = wqc::LQ_SB;
        // The whole factorization is a fixed sequence of launches: its products are staged
        // (one job-table upload) and the panel / T kernels run between them in order.
        struct Act {
            int kind, b0, j0, max_cols;  // kind 0: panel at (b0, j0); 1: T of block b0; 2: extract
        };
        std::vector<std::vector<GemmJob>> calls;
        std::vector<std::vector<Act>> pre(1);  // pre[i]: launches before product i
        auto call = [&](std::vector<GemmJob>& g) {
            calls.push_back(std::move(g));
            pre.emplace_back();
        };
        for (int b0 = 0; b0 < max_n; b0 += SB) {
            // Panels of the block; each updates the block's later rows.
            for (int j0 = b0; j0 < std::min(b0 + SB, max_n); j0 += NB) {
                int max_cols = 1;
                for (const auto& q : lq)
                    if (q.n > j0) max_cols = std::max(max_cols, q.m - b0);
                pre.back().push_back(Act{0, b0, j0, max_cols});
                std::vector<GemmJob> g1, g2;
                for (int i = 0; i < nl; ++i) {
                    const wqc::LqJob& q = lq[i];
                    const int nbp = std::min(NB, q.n - j0);
                    const int nrest = std::min(b0 + SB, q.n) - j0 - nbp;
                    if (nbp <= 0 || nrest <= 0) continue;
                    const int K = q.m - j0;
                    double2* Xr = q.X + (size_t)(j0 + nbp) * q.ldx + j0;
                    g1.push_back(gj(Xr, q.ldx, 0, q.UTc + (size_t)j0 * q.m + j0, q.m, 1, lqw[i], SB, nrest, nbp, K, 1, 0));
                    g2.push_back(gj(lqw[i], SB, 0, q.Ubc + (size_t)j0 * q.m + j0, q.m, 0, Xr, q.ldx, nrest, K, nbp, -1, 1));
                }
                call(g1);
                call(g2);
            }
            // The block's T: Gu = U^H U, then the off-diagonal panels of T.
            std::vector<GemmJob> gg, gt, gh, g1, g2;
            for (int i = 0; i < nl; ++i) {
                const wqc::LqJob& q = lq[i];
                const int nbs = std::min(SB, q.n - b0);
                if (nbs <= 0) continue;
                const int K = q.m - b0;
                const double2* Ub = q.Ubc + (size_t)b0 * q.m + b0;
                gg.push_back(gj(Ub, q.m, 0, Ub, q.m, 1, const_cast<double2*>(q.Gu), SB, nbs, nbs, K, 1, 0));
                gt.push_back(gj(q.T64, SB, 1, Ub, q.m, 0, q.UTc + (size_t)b0 * q.m + b0, q.m, nbs, K, nbs, 1, 0));
                gh.push_back(gj(q.T64, SB, 0, Ub, q.m, 0, q.UThc + (size_t)b0 * q.m + b0, q.m, nbs, K, nbs, 1, 0));
                const int nrest = q.n - b0 - nbs;
                if (nrest <= 0) continue;
                double2* Xr = q.X + (size_t)(b0 + nbs) * q.ldx + b0;
                g1.push_back(gj(Xr, q.ldx, 0, q.UTc + (size_t)b0 * q.m + b0, q.m, 1, lqw[i], SB, nrest, nbs, K, 1, 0));
                g2.push_back(gj(lqw[i], SB, 0, Ub, q.m, 0, Xr, q.ldx, nrest, K, nbs, -1, 1));
            }
            call(gg);
            pre.back().push_back(Act{1, b0, 0, 0});
            // (U T)^H and T U depend only on T: one launch for both.
            gt.insert(gt.end(), gh.begin(), gh.end());
            call(gt);
            // Rows below the block.
            call(g1);
            call(g2);
        }
        pre.back().push_back(Act{2, 0, 0, 0});
        gemm_sequence(calls, 6, [&](size_t i) {
            for (const Act& a : pre[i]) {
                if (a.kind == 0)
                    wqc::launch_lq_panel(dl, nl, a.b0, a.j0, a.max_cols, stream);
                else if (a.kind == 1)
                    wqc::launch_lq_tblock(dl, nl, a.b0, stream);
                else
                    wqc::launch_lq_extract(dl, nl, max_n, stream);
                check_launch();
                if (a.kind != 2) klaunch[6]++;
            }
        });
    }

    // After the Jacobi of LQ-preconditioned jobs: X <- M = [J 0] Q_p^H ... Q_1^H.
    void lq_form_m(const std::vector<wqc::LqJob>& lq, const std::vector<double2*>& lqw) {
        if (lq.empty()) return;
        Timer tm(this, 6);
        const int nl = (int)lq.size();
        wqc::LqJob* dl = upload(lq);
        long mx = 1;
        int max_n = 0;
        for (const auto& q : lq) {
            mx = std::max(mx, (long)q.n * q.m);
            max_n = std::max(max_n, q.n);
        }
        wqc::launch_lq_init_m(dl, nl, mx, stream);
        check_launch();
        const int SB = wqc::LQ_SB;
        std::vector<std::vector<GemmJob>> calls;
        for (int b0 = ((max_n - 1) / SB) * SB; b0 >= 0; b0 -= SB) {
            std::vector<GemmJob> g1, g2;
            for (int i = 0; i < nl; ++i) {
                const wqc::LqJob& q = lq[i];
                const int nbs = std::min(SB, q.n - b0);
                if (nbs <= 0) continue;
                const int K = q.m - b0;
                g1.push_back(gj(q.X + b0, q.ldx, 0, q.UThc + (size_t)b0 * q.m + b0, q.m, 1, lqw[i], SB, q.n, nbs, K, 1, 0));
                g2.push_back(gj(lqw[i], SB, 0, q.Ubc + (size_t)b0 * q.m + b0, q.m, 0, q.X + b0, q.ldx, q.n, K, nbs, -1, 1));
            }
            calls.push_back(std::move(g1));
            calls.push_back(std::move(g2));
        }
        gemm_sequence(calls, 6, [](size_t) {});
    }

    // One Jacobi over the jobs of all batches (range-finder B and exact theta' of a layer
    // share the launches), then sigma and |A|_F^2 of every job read back with one sync.
    void jacobi_splits(std::vector<JobInfo>& J, const std::vector<SplitBatch*>& batches, char* base, int* d_flags) {
        auto PD = [&](size_t o) { return reinterpret_cast<double*>(base + o); };
        // Probe products D = A P of every job, before any preconditioning or rotation touches A.
        size_t njobs = 0, nrows = 0;
        int mmax = 1;
        for (auto* b : batches) {
            b->m0.assign(b->ids.size(), 0);
            for (size_t i = 0; i < b->ids.size(); ++i) {
                b->m0[i] = b->jj[i].m;
                mmax = std::max(mmax, b->jj[i].m);
                nrows += b->jj[i].n;
                ++njobs;
            }
        }
        if (!njobs) return;
        double2* Pm = probe_matrix(mmax);
        probe_buf.ensure(3 * nrows * 2 * sizeof(double2) + 2 * njobs * sizeof(double) + 256);
        note_peak();
        double2* pd = reinterpret_cast<double2*>(probe_buf.p);
        double2* pt = pd + nrows * 2;
        double2* pr = pt + nrows * 2;
        double* perr = reinterpret_cast<double*>(pr + nrows * 2);
        std::vector<size_t> prow;  // first row of each job in the probe buffers
        {
            Timer tm(this, 4);
            std::vector<GemmJob> g;
            size_t at = 0;
            for (auto* b : batches)
                for (size_t i = 0; i < b->ids.size(); ++i) {
                    const JacobiJob& x = b->jj[i];
                    prow.push_back(at);
                    g.push_back(gj(x.X, x.ldx, 0, Pm, 2, 0, pd + 2 * at, 2, x.n, 2, x.m, 1, 0));
                    at += x.n;
                }
            gemm(g, 4);
        }
        std::vector<wqc::LqJob> lq, lq2;
        std::vector<double2*> lqw, lqw2;
        lq_precondition(J, batches, base, lq, lqw, lq2, lqw2);
        std::vector<JacobiJob> all;
        std::vector<long> offs;
        for (auto* b : batches) {
            all.insert(all.end(), b->jj.begin(), b->jj.end());
            offs.insert(offs.end(), b->offs.begin(), b->offs.end());
        }
        if (all.empty()) return;
        jacobi_any(all, PD(0), offs, d_flags);
        if (lq2.empty()) {
            lq_form_m(lq, lqw);
        } else {
            // X = L Q = R^H Q and R = L2 Q2, so X = (Q2^H J^H)(Y2 Q) with Y2 = J R2 the rotated rows:
            // the plain form's X = J_eff^H X_eff with J_eff = J Q2 (into R's buffer) and
            // X_eff = [Y2 0] Q (into X's own buffer), both by the reflectors applied in reverse.
            lq_form_m(lq2, lqw2);
            for (size_t k = 0; k < lq.size(); ++k) lq[k].J = lq2[k].R;
            lq_form_m(lq, lqw);
            size_t k = 0;
            for (auto* b : batches)
                for (size_t i = 0; i < b->ids.size(); ++i) {
                    if (!b->pre[i]) continue;
                    JacobiJob& x = b->jj[i];
                    x.J = lq2[k].X;
                    x.X = b->orig[i];
                    x.ldx = b->orig_ld[i];
                    x.m = b->m0[i];
                    b->pre[i] = 0;
                    ++k;
                }
        }
        // Rebuilt probes U S V^H x = F^H (G x): (F, G) = (J, rotated rows) for a plain job
        // (A = J^H X_final), (rotated rows of R, M) for an LQ-preconditioned one (A = Y^H M).
        {
            Timer tm(this, 4);
            std::vector<GemmJob> g1, g2;
            std::vector<wqc::ProbeJob> pj;
            size_t t = 0;
            for (auto* b : batches)
                for (size_t i = 0; i < b->ids.size(); ++i, ++t) {
                    const JacobiJob& x = b->jj[i];
                    const bool pre = !b->pre.empty() && b->pre[i];
                    const double2* F = pre ? x.X : x.J;
                    const int ldf = pre ? x.ldx : x.n;
                    const double2* G = pre ? b->orig[i] : x.X;
                    const int ldg = pre ? b->orig_ld[i] : x.ldx;
                    double2* tt = pt + 2 * prow[t];
                    double2* rr = pr + 2 * prow[t];
                    g1.push_back(gj(G, ldg, 0, Pm, 2, 0, tt, 2, x.n, 2, b->m0[i], 1, 0));
                    g2.push_back(gj(F, ldf, 1, tt, 2, 0, rr, 2, x.n, 2, x.n, 1, 0));
                    pj.push_back(wqc::ProbeJob{pd + 2 * prow[t], rr, perr + 2 * t, x.n, 0});
                }
            gemm(g1, 4);
            gemm(g2, 4);
            wqc::launch_probe_residual(upload(pj), (int)pj.size(), stream);
            check_launch();
        }
        if (trace_jacobi) {
            std::fprintf(stderr, "jacobi splits (site:rows x cols:sweeps):");
            size_t t = 0;
            for (auto* b : batches)
                for (size_t i = 0; i < b->ids.size(); ++i, ++t)
                    std::fprintf(stderr, " %d:%dx%d:%d", J[b->ids[i]].p, b->jj[i].n, b->jj[i].m, last_sweeps[t]);
            std::fprintf(stderr, "\n");
        }
        // sigma of every job and |A|_F^2 of every batch, gathered on the device into one
        // contiguous block and read back with one copy.
        size_t total = 2 * njobs;
        for (auto* b : batches) {
            for (size_t i = 0; i < b->ids.size(); ++i) total += b->jj[i].n;
            total += b->ids.size();
        }
        if (total + 1 > hread_cap) fail(WQC_TN_EXEC_FAILED, "readback buffer too small");
        gather.ensure(std::max<size_t>(total, 1) * sizeof(double));
        note_peak();
        double* gd = reinterpret_cast<double*>(gather.p);
        std::vector<const double*> cs;
        std::vector<double*> cd;
        std::vector<long> cc;
        size_t k = 0, cur = 0;
        for (auto* b : batches) {
            b->hoff.assign(b->ids.size(), 0);
            for (size_t i = 0; i < b->ids.size(); ++i, ++k) {
                J[b->ids[i]].sweeps += last_sweeps[k];
                J[b->ids[i]].sweeps_lo += last_sweeps_lo[k];
                b->hoff[i] = cur;
                cs.push_back(PD(0) + b->offs[i]);
                cd.push_back(gd + cur);
                cc.push_back(b->jj[i].n);
                cur += b->jj[i].n;
            }
        }
        for (auto* b : batches) {
            if (b->ids.empty()) continue;
            b->hfro = cur;
            cs.push_back(b->d_fro);
            cd.push_back(gd + cur);
            cc.push_back((long)b->ids.size());
            cur += b->ids.size();
        }
        // The probe residuals of every job, in batch order.
        {
            size_t t0 = 0;
            for (auto* b : batches) {
                b->hprobe = cur + 2 * t0;
                t0 += b->ids.size();
            }
            cs.push_back(perr);
            cd.push_back(gd + cur);
            cc.push_back((long)(2 * njobs));
            cur += 2 * njobs;
        }
        copy_batch(cs, cd, cc);
        if (cur) HIPC(hipMemcpyAsync(hread, gd, cur * sizeof(double), hipMemcpyDeviceToHost, stream));
        sync();
    }

    // Exact split, after the Jacobi: stable descending sigma, the norm check of
    // svd_reproduces, retained_chi, reject_truncation, sqrt(sigma) factors.
    void finish_exact(std::vector<JobInfo>& J, const SplitBatch& b, char* base) {
        auto PD = [&](size_t o) { return reinterpret_cast<double*>(base + o); };
        const std::vector<int>& ids = b.ids;
        const std::vector<JacobiJob>& jj = b.jj;
        const int ne = (int)ids.size();
        std::vector<FactorJob> fj(ne);
        std::vector<std::vector<int>> perms(ne);
        long mxl = 1, mxr = 1;
        for (int k = 0; k < ne; ++k) {
            JobInfo& j = J[ids[k]];
            const int nn = jj[k].n;
            std::vector<double> sig(hread + b.hoff[k], hread + b.hoff[k] + nn);
            const double fro2 = hread[b.hfro + k];
            std::vector<int> order(nn);
            for (int i = 0; i < nn; ++i) order[i] = i;
            std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return sig[a] > sig[b]; });
            std::vector<double> sorted(nn);
            double s2 = 0.0;
            for (int i = 0; i < nn; ++i) {
                sorted[i] = sig[order[i]];
                s2 += sorted[i] * sorted[i];
            }
            const bool norm_ok =
                std::fabs(s2 - fro2) <= SVD_CHECK_TOL * std::max(fro2, 2.2250738585072014e-308) + 1e-28;
            j.probe = probe_rel(b, k, fro2);
            if (!norm_ok || !probes_pass(b, k, fro2)) {
                j.st->stats.svd_check_failures++;
                char buf[256];
                std::snprintf(buf, sizeof buf, " (sum s^2 = %.17g, |A|_F^2 = %.17g, probe residuals %.3g %.3g, job %d of %d, n=%d m=%d)",
                              s2, fro2, std::sqrt(hread[b.hprobe + 2 * (size_t)k]),
                              std::sqrt(hread[b.hprobe + 2 * (size_t)k + 1]), k, ne, nn, b.m0[k]);
                fail(WQC_TN_EXEC_FAILED, "SVD of a " + std::to_string(j.rows) + "x" + std::to_string(j.cols) +
                                             " matrix failed its check" + buf);
            }
            const size_t c = retained_chi(sorted, j.st->chi, j.rows, j.cols);
            if (trace_jacobi)
                for (int i = 0; i < nn; ++i)
                    if (sorted[i] > 0.2 * SINGULAR_CUTOFF && sorted[i] < 5 * SINGULAR_CUTOFF)
                        std::fprintf(stderr, "near cutoff: site %d %dx%d kept %zu sigma[%d] = %.6e\n", j.p, j.rows,
                                     j.cols, c, i, sorted[i]);
            bool dropped = false;
            double drop = 0.0;
            for (int i = (int)c; i < nn; ++i) {
                if (sorted[i] > SINGULAR_CUTOFF) dropped = true;
                drop += sorted[i] * sorted[i];
            }
            wqc_tn_stats& sx = j.st->stats;
            if (j.st->reject && dropped) fail(WQC_TN_BOND_EXCEEDED, "mps bond exceeded chi");
            if (dropped) sx.truncated_splits++;
            const double dw = s2 > 0 ? drop / s2 : 0.0;
            j.dw = dw;
            sx.max_discarded_weight = std::max(sx.max_discarded_weight, dw);
            sx.sum_discarded_weight += dw;
            sx.exact_splits++;
            perms[k].assign(order.begin(), order.begin() + c);
            j.chi_kept = c;
            grow_site(j.st, j.p, (size_t)j.dl * 2 * c);
            grow_site(j.st, j.p + 1, c * 2 * (size_t)j.dr);
            const bool tall = j.rows > j.cols;
            const bool pre = !b.pre.empty() && b.pre[k];
            fj[k].X = jj[k].X;
            fj[k].J = pre ? b.orig[k] : jj[k].J;
            fj[k].left = j.st->sites[j.p].ptr;
            fj[k].right = j.st->sites[j.p + 1].ptr;
            fj[k].perm = nullptr;  // set below from the batch's one upload of every perm
            fj[k].sigma = PD(j.o_sig);
            fj[k].rows = j.rows;
            fj[k].cols = j.cols;
            fj[k].chi = (int)c;
            fj[k].ldx = jj[k].ldx;
            fj[k].ldj = pre ? b.orig_ld[k] : nn;
            fj[k].mode = (tall != pre) ? 1 : 0;
            mxl = std::max(mxl, (long)j.rows * (long)c);
            mxr = std::max(mxr, (long)c * j.cols);
        }
        if (tape)
            for (int k = 0; k < ne; ++k) {
                JobInfo& j = J[ids[k]];
                if (j.rec < 0) continue;
                TapeUnit& t = tape->units[j.rec];
                const int nn = jj[k].n, mm = jj[k].m;
                t.pp = nn;
                t.qq = mm;
                t.c = (int)j.chi_kept;
                t.tall = j.rows > j.cols;
                t.krow = perms[k];
                t.sig.assign(hread + b.hoff[k], hread + b.hoff[k] + nn);
                t.X = tape->arena.take((size_t)nn * mm);
                t.J = tape->arena.take((size_t)nn * nn);
                HIPC(hipMemcpy2DAsync(t.X, (size_t)mm * sizeof(double2), jj[k].X, (size_t)jj[k].ldx * sizeof(double2),
                                      (size_t)mm * sizeof(double2), nn, hipMemcpyDeviceToDevice, stream));
                HIPC(hipMemcpyAsync(t.J, jj[k].J, (size_t)nn * nn * sizeof(double2), hipMemcpyDeviceToDevice, stream));
                t.done = true;
            }
        set_perms(fj, perms);
        {
            Timer tm(this, 4);
            wqc::launch_factor(upload(fj), ne, mxl, mxr, stream);
            check_launch();
            klaunch[4]++;
        }
        for (int k = 0; k < ne; ++k) {
            JobInfo& j = J[ids[k]];
            j.st->sites[j.p].r = (int)j.chi_kept;
            j.st->sites[j.p + 1].l = (int)j.chi_kept;
        }
    }

    // Range finder, after the Jacobi of B: dust on sigma, retained_chi, needs_f64_fallback
    // (fallback jobs are marked for an exact split), factors, left = Q W.
    void finish_rsvd(std::vector<JobInfo>& J, const SplitBatch& b, char* base) {
        auto P = [&](size_t o) { return reinterpret_cast<double2*>(base + o); };
        auto PD = [&](size_t o) { return reinterpret_cast<double*>(base + o); };
        const std::vector<int>& ids = b.ids;
        const std::vector<JacobiJob>& jj = b.jj;
        const int nr = (int)ids.size();
        std::vector<FactorJob> fj;
        std::vector<int> kept;
        std::vector<int> kept_k;  // batch index k of each kept job
        std::vector<std::vector<int>> perms(nr);
        long mxl = 1, mxr = 1;
        for (int k = 0; k < nr; ++k) {
            JobInfo& j = J[ids[k]];
            const int nn = jj[k].n;
            std::vector<double> sig(hread + b.hoff[k], hread + b.hoff[k] + nn);
            const double fro2 = hread[b.hfro + k];
            std::vector<int> order(nn);
            for (int i = 0; i < nn; ++i) order[i] = i;
            std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return sig[a] > sig[b]; });
            std::vector<double> values(nn);
            double s2 = 0.0;
            for (int i = 0; i < nn; ++i) {
                values[i] = sig[order[i]];
                s2 += values[i] * values[i];
            }
            const bool svd_ok = std::fabs(s2 - fro2) <= SVD_CHECK_TOL * std::max(fro2, 2.2250738585072014e-308) + 1e-28 &&
                                probes_pass(b, k, fro2);
            j.probe = probe_rel(b, k, fro2);
            wqc_tn_stats& sx = j.st->stats;
            if (!svd_ok) sx.svd_check_failures++;
            for (double& s : values)
                if (s < SINGULAR_CUTOFF * 1e-2) s = 0.0;
            const size_t c = retained_chi(values, j.st->chi, j.rows, j.cols);
            if (trace_jacobi && (!svd_ok || c == 0 || needs_f64_fallback(values, c))) {
                std::fprintf(stderr, "rsvd fallback: site %d ell %zu kept %zu svd_ok %d", j.p, j.ell, c, (int)svd_ok);
                for (int i = 0; i < nn; ++i)
                    if (values[i] > 0.0 && std::fabs(values[i] - SINGULAR_CUTOFF) <= CUTOFF_ABS_BAND)
                        std::fprintf(stderr, " s[%d]=%.3e", i, values[i]);
                if (c > 0 && c < (size_t)nn) std::fprintf(stderr, " kept %.6e dropped %.6e", values[c - 1], values[c]);
                std::fprintf(stderr, "\n");
            }
            if (!svd_ok || c == 0 || needs_f64_fallback(values, c)) {
                j.fallback = true;
                sx.rsvd_fallbacks++;
                continue;
            }
            sx.rsvd_kept++;
            bool dropped = false;
            double drop = 0.0, tot = 0.0;
            for (int i = 0; i < nn; ++i) {
                tot += values[i] * values[i];
                if (i >= (int)c) {
                    drop += values[i] * values[i];
                    if (values[i] > SINGULAR_CUTOFF) dropped = true;
                }
            }
            if (dropped) sx.truncated_splits++;
            const double dw = tot > 0 ? drop / tot : 0.0;
            j.dw = dw;
            sx.max_discarded_weight = std::max(sx.max_discarded_weight, dw);
            sx.sum_discarded_weight += dw;
            perms[k].assign(order.begin(), order.begin() + c);
            j.chi_kept = c;
            grow_site(j.st, j.p, (size_t)j.dl * 2 * c);
            grow_site(j.st, j.p + 1, c * 2 * (size_t)j.dr);
            FactorJob f;
            const bool pre = !b.pre.empty() && b.pre[k];
            f.X = jj[k].X;
            f.J = pre ? b.orig[k] : jj[k].J;
            f.left = P(j.o_W);
            f.right = j.st->sites[j.p + 1].ptr;
            f.perm = nullptr;  // set below from the batch's one upload of every perm
            f.sigma = PD(j.o_sigB);
            f.rows = j.rows;
            f.cols = j.cols;
            f.chi = (int)c;
            f.ldx = jj[k].ldx;
            f.ldj = pre ? b.orig_ld[k] : (int)j.ell;
            f.mode = pre ? 3 : 2;
            f.pad0 = f.pad1 = 0;
            fj.push_back(f);
            kept.push_back(ids[k]);
            kept_k.push_back(k);
            mxl = std::max(mxl, (long)j.ell * (long)c);
            mxr = std::max(mxr, (long)c * j.cols);
        }
        if (fj.empty()) return;
        {
            std::vector<std::vector<int>> kp;
            for (size_t q = 0; q < kept.size(); ++q) kp.push_back(std::move(perms[kept_k[q]]));
            set_perms(fj, kp);
        }
        Timer tm(this, 4);
        wqc::launch_factor(upload(fj), (int)fj.size(), mxl, mxr, stream);
        check_launch();
        klaunch[4]++;
        // left = [W; 0] - V (T (V[0:ell, :]^H W))
        std::vector<GemmJob> g7, g8, g9;
        for (int i : kept) {
            JobInfo& j = J[i];
            const int r = j.om->r, c = (int)j.chi_kept;
            double2* W2 = P(j.o_W2);
            double2* W3 = W2 + (size_t)r * c;
            g7.push_back(gj(P(j.o_V), r, 1, P(j.o_W), c, 0, W2, c, r, c, (int)j.ell, 1, 0));
            g8.push_back(gj(P(j.o_T), r, 0, W2, c, 0, W3, c, r, c, r, 1, 0));
            double2* left = j.st->sites[j.p].ptr;
            HIPC(hipMemcpyAsync(left, P(j.o_W), j.ell * (size_t)c * sizeof(double2), hipMemcpyDeviceToDevice, stream));
            if ((size_t)j.rows > j.ell)
                HIPC(hipMemsetAsync(left + j.ell * (size_t)c, 0, ((size_t)j.rows - j.ell) * c * sizeof(double2), stream));
            g9.push_back(gj(P(j.o_V), r, 0, W3, c, 0, left, c, j.rows, c, r, -1, 1));
        }
        gemm(g7);
        gemm(g8);
        gemm(g9);
        for (int i : kept) {
            JobInfo& j = J[i];
            j.st->sites[j.p].r = (int)j.chi_kept;
            j.st->sites[j.p + 1].l = (int)j.chi_kept;
        }
    }

    // ------------------------------------------------------------ scheduling
    // With `defer` set, flush() records the pending site-disjoint operations as one layer
    // instead of running them: the slice batch runs the layers of many states together.
    bool defer = false;
    struct Layer {
        std::vector<PendingOne> ones;
        std::vector<Unit> units;
    };
    std::vector<Layer> layers;

    void flush() {
        if (defer) {
            if (!pend1.empty() || !pend2.empty()) {
                layers.push_back(Layer{std::move(pend1), std::move(pend2)});
                pend1.clear();
                pend2.clear();
            }
            std::fill(touched.begin(), touched.end(), 0);
            return;
        }
        if (!pend1.empty()) {
            std::vector<PendingOne> ops;
            ops.swap(pend1);
            run_one_qubit(ops);
        }
        if (!pend2.empty()) {
            std::vector<Unit> units;
            units.swap(pend2);
            run_units(units);
        }
        std::fill(touched.begin(), touched.end(), 0);
    }

    void queue_one(int site, const Mat2 u) {
        if (touched[site]) flush();
        PendingOne op;
        op.site = site;
        for (int k = 0; k < 4; ++k) op.u[k] = u[k];
        pend1.push_back(op);
        touched[site] = 1;
    }

    void queue_unit(int p, const Mat4 u, bool ctrl_is_first) {
        if (touched[p] || touched[p + 1]) flush();
        check_cancel();
        Unit un;
        un.p = p;
        resolve_u4(u, ctrl_is_first, un.u4);
        pend2.push_back(un);
        touched[p] = touched[p + 1] = 1;
    }

    // apply_two_qubit_gate (src/tn/mps.rs:353-375): bubble `right` down to left+1 with
    // SWAP units, apply, bubble back.
    void queue_two(int ctrl, int tgt, const Mat4 u) {
        const int left = std::min(ctrl, tgt), right = std::max(ctrl, tgt);
        const int span = right - left;
        Mat4 sw;
        two_qubit_matrix(KIND_SWAP, sw);
        if (span > 1)
            for (int pos = right - 1; pos >= left + 1; --pos) queue_unit(pos, sw, true);
        queue_unit(left, u, ctrl < tgt);
        if (span > 1)
            for (int pos = left + 1; pos < right; ++pos) queue_unit(pos, sw, true);
    }

    void check_gate(const wqc_tn_gate& g) {
        auto bad = [&](int q) { return q < 0 || q >= n; };
        const int k = g.kind;
        if (is_one_qubit(k) || k == WQC_TN_GATE_MEASURE || k == WQC_TN_GATE_RESET) {
            if (bad(g.q0)) fail(WQC_TN_INVALID_ARG, "qubit index " + std::to_string(g.q0) + " out of bounds (limit: " + std::to_string(n) + ")");
        } else if (k == WQC_TN_GATE_CNOT || k == WQC_TN_GATE_CZ) {
            if (bad(g.q0) || bad(g.q1)) fail(WQC_TN_INVALID_ARG, "qubit index out of bounds");
            if (g.q0 == g.q1) fail(WQC_TN_INVALID_ARG, "two-qubit gate on one qubit");
        } else if (k == WQC_TN_GATE_CCNOT) {
            if (bad(g.q0) || bad(g.q1) || bad(g.q2)) fail(WQC_TN_INVALID_ARG, "qubit index out of bounds");
            if (g.q0 == g.q1 || g.q0 == g.q2 || g.q1 == g.q2) fail(WQC_TN_INVALID_ARG, "CCNOT needs three distinct qubits");
        } else if (k != WQC_TN_GATE_IF) {
            fail(WQC_TN_INVALID_ARG, "unknown gate kind " + std::to_string(k));
        }
    }

    // apply_gate (src/tn/mps.rs:278-297), queued.
    void queue_gate(const wqc_tn_gate& g) {
        check_gate(g);
        const int k = g.kind;
        if (is_one_qubit(k)) {
            Mat2 u;
            unary_matrix(k, g.angle, u);
            queue_one(g.q0, u);
        } else if (k == WQC_TN_GATE_CNOT || k == WQC_TN_GATE_CZ) {
            Mat4 u;
            two_qubit_matrix(k, u);
            queue_two(g.q0, g.q1, u);
        } else if (k == WQC_TN_GATE_CCNOT) {
            queue_ccnot(g.q0, g.q1, g.q2);
        }
    }

    // apply_ccnot (src/tn/mps.rs:657-677).
    void queue_ccnot(int c1, int c2, int t) {
        Mat2 h, tg, tdg;
        unary_matrix(WQC_TN_GATE_H, 0, h);
        unary_matrix(WQC_TN_GATE_T, 0, tg);
        unary_matrix(0, 0, tdg);
        Mat4 cn;
        two_qubit_matrix(WQC_TN_GATE_CNOT, cn);
        queue_one(t, h);
        queue_two(c2, t, cn);
        queue_one(t, tdg);
        queue_two(c1, t, cn);
        queue_one(t, tg);
        queue_two(c2, t, cn);
        queue_one(t, tdg);
        queue_two(c1, t, cn);
        queue_one(t, tg);
        queue_one(t, h);
        queue_one(c2, tg);
        queue_two(c1, c2, cn);
        queue_one(c1, tg);
        queue_one(c2, tdg);
        queue_two(c1, c2, cn);
    }

    // The apply loop of execute_with_trace_using (src/tn/trace.rs:71-161).
    void queue_circuit(const wqc_tn_gate* gates, size_t count) {
        size_t i = 0;
        while (i < count) {
            const wqc_tn_gate& g = gates[i];
            check_gate(g);
            if (g.kind == WQC_TN_GATE_MEASURE || g.kind == WQC_TN_GATE_RESET || g.kind == WQC_TN_GATE_IF) {
                ++i;
                continue;
            }
            if (g.kind == WQC_TN_GATE_RX || g.kind == WQC_TN_GATE_RY || g.kind == WQC_TN_GATE_RZ) {
                size_t run = 1;
                double sum = g.angle;
                while (i + run < count && gates[i + run].kind == g.kind && gates[i + run].q0 == g.q0) {
                    sum += gates[i + run].angle;
                    run++;
                }
                const double tau = 6.283185307179586;
                double red = std::fmod(sum, tau);
                if (red < 0) red += tau;
                const bool identity = std::fabs(red) < 1e-9 || std::fabs(tau - red) < 1e-9;
                if (identity || run == 1) {
                    for (size_t k = 0; k < run; ++k) queue_gate(gates[i + k]);
                } else {
                    wqc_tn_gate net = g;
                    net.angle = sum;
                    queue_gate(net);
                }
                i += run;
                continue;
            }
            queue_gate(g);
            ++i;
        }
    }

    // ------------------------------------------------------------ overlap
    // <bra|ket> with this state as the ket (overlap, src/tn/mps.rs:1421-1474): the environment
    // (ket bond x bra bond) is carried left to right and may hold at most 256 entries.
    DevBuf ov_buf;
    void overlap_with(wqc_tn_state& bra, double out[2]) {
        if (bra.n != n) fail(WQC_TN_INVALID_ARG, "mismatched register");
        flush();
        bra.flush();
        bra.sync();
        ov_buf.ensure(2 * wqc::OVERLAP_ENV_MAX * sizeof(double2));
        double2* cur = reinterpret_cast<double2*>(ov_buf.p);
        double2* nxt = cur + wqc::OVERLAP_ENV_MAX;
        HIPC(hipMemcpyAsync(cur, &kOne, sizeof(double2), hipMemcpyHostToDevice, stream));
        int kl = 1, bl = 1;
        for (int i = 0; i < n; ++i) {
            const Site& K = sites[i];
            const Site& B = bra.sites[i];
            if (K.l != kl || B.l != bl) fail(WQC_TN_EXEC_FAILED, "mps site shape does not match the overlap");
            if ((long)K.r * B.r > wqc::OVERLAP_ENV_MAX) fail(WQC_TN_EXEC_FAILED, "mps overlap environment exceeds 256");
            wqc::launch_overlap_site(cur, K.ptr, B.ptr, nxt, kl, K.r, bl, B.r, stream);
            check_launch();
            std::swap(cur, nxt);
            kl = K.r;
            bl = B.r;
        }
        if (kl * bl != 1) fail(WQC_TN_EXEC_FAILED, "mps overlap left an open bond");
        double2 v;
        HIPC(hipMemcpyAsync(&v, cur, sizeof(double2), hipMemcpyDeviceToHost, stream));
        sync();
        out[0] = v.x;
        out[1] = v.y;
    }

    // ------------------------------------------------------------ readouts
    // Contract a vector through sites [from, to) (left to right) or (to, from] (right to
    // left) with physical weight phys (0, 1 or 2 = sum). Returns the device vector.
    double2* env_pass(int from, int to, int phys, bool right_to_left, int& len, double2* bufA,
                      double2* bufB) {
        HIPC(hipMemcpyAsync(bufA, &kOne, sizeof(double2), hipMemcpyHostToDevice, stream));
        len = 1;
        double2* cur = bufA;
        double2* nxt = bufB;
        std::vector<VecMatJob> jobs;
        std::vector<int> outl, inl;
        if (!right_to_left) {
            int vlen = 1;
            for (int i = from; i < to; ++i) {
                VecMatJob v{sites[i].ptr, cur, nxt, sites[i].l, sites[i].r, vlen, phys, 0, 0, vm_part};
                jobs.push_back(v);
                outl.push_back(sites[i].r);
                inl.push_back(std::min(sites[i].l, vlen));
                vlen = sites[i].r;
                std::swap(cur, nxt);
            }
        } else {
            int vlen = 1;
            for (int i = from; i > to; --i) {
                VecMatJob v{sites[i].ptr, cur, nxt, sites[i].l, sites[i].r, vlen, phys, 1, 0, vm_part};
                jobs.push_back(v);
                outl.push_back(sites[i].l);
                inl.push_back(std::min(sites[i].r, vlen));
                vlen = sites[i].l;
                std::swap(cur, nxt);
            }
        }
        if (!jobs.empty()) {
            VecMatJob* d = upload(jobs);
            for (size_t k = 0; k < jobs.size(); ++k)
                wqc::launch_vecmat(d + k, outl[k], inl[k], right_to_left ? 1 : 0, stream);
            check_launch();
            len = outl.back();
        }
        return cur;
    }
    static constexpr double2 kOne = {1.0, 0.0};

    // Four readout vectors and the left pass's chunk partials (vm_part).
    double2* vm_part = nullptr;
    void readout_buffers(double2*& a, double2*& b, double2*& c, double2*& d) {
        size_t mx = 2;
        for (auto& s : sites) mx = std::max(mx, (size_t)std::max(s.l, s.r) + 2);
        const size_t part = (mx / wqc::kVecMatChunk + 2) * mx;
        ws.ensure((4 * mx + part) * sizeof(double2) + 4096);
        note_peak();
        double2* base = (double2*)ws.p;
        a = base;
        b = base + mx;
        c = base + 2 * mx;
        d = base + 3 * mx;
        vm_part = base + 4 * mx;
    }

    std::pair<double, double> amplitude_zero() {
        flush();
        sync();
        tables_reset();
        klaunch[5] += n;
        Timer tm(this, 5);
        double2 *a, *b, *c, *d;
        readout_buffers(a, b, c, d);
        int len = 0;
        double2* v = env_pass(0, n, 0, false, len, a, b);
        double2 out;
        HIPC(hipMemcpyAsync(&out, v, sizeof(double2), hipMemcpyDeviceToHost, stream));
        sync();
        (void)c;
        (void)d;
        return {out.x, out.y};
    }
    // amplitude_zero with the chain vectors and partials in the device's scratch array instead of
    // the state's workspace: nothing is allocated (no device buffer grows, no event is created).
    std::pair<double, double> amplitude_zero_noalloc() {
        if (!pend1.empty() || !pend2.empty()) fail(WQC_TN_EXEC_FAILED, "state has unapplied gates");
        double2* base = dev->scratch;
        if (!base) fail(WQC_TN_EXEC_FAILED, "device scratch not set up");
        size_t mx = 2;
        for (auto& s : sites) mx = std::max(mx, (size_t)std::max(s.l, s.r) + 2);
        const size_t need = 2 * mx + (mx / wqc::kVecMatChunk + 2) * mx;
        if (need > kScratchElems)
            fail(WQC_TN_INVALID_ARG, "bond dimension " + std::to_string(mx - 2) +
                                         " too large for the allocation-free readout: it needs " +
                                         std::to_string(need * sizeof(double2)) + " bytes of scratch, the array holds " +
                                         std::to_string(kScratchElems * sizeof(double2)) + " (bonds up to 1024 fit)");
        sync();
        tables_reset();
        double2* saved = vm_part;
        vm_part = base + 2 * mx;
        int len = 0;
        double2* v = env_pass(0, n, 0, false, len, base, base + mx);
        vm_part = saved;
        klaunch[5] += n;
        double2 out;
        HIPC(hipMemcpyAsync(&out, v, sizeof(double2), hipMemcpyDeviceToHost, stream));
        sync();
        return {out.x, out.y};
    }

    // <0...0|psi> of another state's sites (a slice container of the batch), computed with this
    // state's stream and buffers.
    // <0...0|psi> of many states together: position q of every state's vector-times-site chain
    // runs in one launch pair (each state's vector ping-pongs in its own two buffers), so a
    // batch of slices costs one launch pair per site position instead of one per site.
    DevBuf amp_buf;
    void amplitudes_batch(const std::vector<wqc_tn_state*>& states, double* out) {
        flush();
        sync();
        tables_reset();
        const size_t ns = states.size();
        size_t mx = 2;
        int maxn = 0;
        for (auto* s : states) {
            if (!s) continue;
            maxn = std::max(maxn, s->n);
            for (const auto& x : s->sites) mx = std::max(mx, (size_t)std::max(x.l, x.r) + 2);
        }
        const size_t part = (mx / wqc::kVecMatChunk + 2) * mx;
        const size_t per = 2 * mx + part;
        amp_buf.ensure(std::max<size_t>(ns, 1) * per * sizeof(double2) + 4096);
        note_peak();
        double2* base = (double2*)amp_buf.p;
        std::vector<double2> ones(ns, kOne);
        HIPC(hipMemcpy2DAsync(base, per * sizeof(double2), ones.data(), sizeof(double2), sizeof(double2), ns,
                              hipMemcpyHostToDevice, stream));
        std::vector<int> vlen(ns, 1), cur(ns, 0);
        Timer tm(this, 5);
        for (int q = 0; q < maxn; ++q) {
            std::vector<VecMatJob> jobs;
            int mo = 1, mi = 1;
            for (size_t i = 0; i < ns; ++i) {
                if (!states[i] || q >= states[i]->n) continue;
                const Site& x = states[i]->sites[q];
                double2* b = base + i * per;
                jobs.push_back(VecMatJob{x.ptr, b + cur[i] * mx, b + (1 - cur[i]) * mx, x.l, x.r, vlen[i], 0, 0, 0,
                                         b + 2 * mx});
                mo = std::max(mo, x.r);
                mi = std::max(mi, std::min(x.l, vlen[i]));
                vlen[i] = x.r;
                cur[i] ^= 1;
            }
            wqc::launch_vecmat_batch(upload(jobs), (int)jobs.size(), mo, mi, stream);
            check_launch();
            klaunch[5]++;
        }
        // Every chain ends in a vector of length 1: element 0 of the buffer it was last written to.
        std::vector<double2> h(2 * ns);
        HIPC(hipMemcpy2DAsync(h.data(), 2 * sizeof(double2), base, per * sizeof(double2), sizeof(double2), ns,
                              hipMemcpyDeviceToHost, stream));
        HIPC(hipMemcpy2DAsync(h.data() + 1, 2 * sizeof(double2), base + mx, per * sizeof(double2), sizeof(double2),
                              ns, hipMemcpyDeviceToHost, stream));
        sync();
        for (size_t i = 0; i < ns; ++i) {
            if (!states[i]) continue;
            const double2 v = h[2 * i + cur[i]];
            out[2 * i] = v.x;
            out[2 * i + 1] = v.y;
        }
    }

    std::pair<double, double> amplitude_zero_timed() {
        auto r = amplitude_zero();
        resolve_timings();
        return r;
    }

    // site_amplitudes (src/tn/mps.rs:679-693): left/right environments with the all-ones
    // physical vector, then a_s = sum_{l,r} le[l] site[l,s,r] re[r].
    void site_amplitudes(int t, double out[4]) {
        flush();
        sync();
        tables_reset();
        double2 *a, *b, *c, *d;
        readout_buffers(a, b, c, d);
        int llen = 0, rlen = 0;
        double2* le = env_pass(0, t, 2, false, llen, a, b);
        double2* re = env_pass(n - 1, t, 2, true, rlen, c, d);
        std::vector<VecMatJob> jobs;
        double2* scratch0 = (le == a) ? b : a;
        // u_s = le . site[:, s, :]  (length site.r), then a_s = u_s . re
        double2 *u0 = scratch0, *u1 = nullptr;
        double2* freec = (re == c) ? d : c;
        u1 = freec;
        VecMatJob j0{sites[t].ptr, le, u0, sites[t].l, sites[t].r, llen, 0, 0, 0, vm_part};
        VecMatJob j1{sites[t].ptr, le, u1, sites[t].l, sites[t].r, llen, 1, 0, 0, vm_part};
        jobs.push_back(j0);
        jobs.push_back(j1);
        VecMatJob* dj = upload(jobs);
        wqc::launch_vecmat(dj, sites[t].r, std::min(sites[t].l, llen), 0, stream);
        wqc::launch_vecmat(dj + 1, sites[t].r, std::min(sites[t].l, llen), 0, stream);
        check_launch();
        const int rr = sites[t].r;
        std::vector<double2> hu0(rr), hu1(rr), hre(std::max(rlen, 1));
        HIPC(hipMemcpyAsync(hu0.data(), u0, rr * sizeof(double2), hipMemcpyDeviceToHost, stream));
        HIPC(hipMemcpyAsync(hu1.data(), u1, rr * sizeof(double2), hipMemcpyDeviceToHost, stream));
        HIPC(hipMemcpyAsync(hre.data(), re, rlen * sizeof(double2), hipMemcpyDeviceToHost, stream));
        sync();
        double2 a0 = make_double2(0, 0), a1 = make_double2(0, 0);
        for (int r = 0; r < rr; ++r) {
            const double2 w = r < rlen ? hre[r] : make_double2(1, 0);
            a0.x += hu0[r].x * w.x - hu0[r].y * w.y;
            a0.y += hu0[r].x * w.y + hu0[r].y * w.x;
            a1.x += hu1[r].x * w.x - hu1[r].y * w.y;
            a1.y += hu1[r].x * w.y + hu1[r].y * w.x;
        }
        out[0] = a0.x;
        out[1] = a0.y;
        out[2] = a1.x;
        out[3] = a1.y;
    }

    // site_amplitudes for many targets of one state (the trace's first-touch probes and control
    // probabilities of a layer, src/tn/trace.rs:346-420): the all-ones left environments of
    // every prefix and right environments of every suffix in one sweep each, then one launch
    // for all targets; out[4k..4k+3] = (a0, a1) of targets[k].
    void site_amplitudes_batch(const int32_t* targets, size_t count, double* out) {
        flush();
        sync();
        tables_reset();
        Timer tm(this, 5);
        std::vector<size_t> loff(n + 1), roff(n + 2);
        std::vector<int> llen(n + 1), rlen(n + 2);
        size_t total = 0;
        llen[0] = 1;
        loff[0] = total;
        total += 1;
        for (int q = 0; q < n; ++q) {
            llen[q + 1] = sites[q].r;
            loff[q + 1] = total;
            total += sites[q].r;
        }
        rlen[n] = 1;
        roff[n] = total;
        total += 1;
        for (int q = n - 1; q >= 0; --q) {
            rlen[q] = sites[q].l;
            roff[q] = total;
            total += sites[q].l;
        }
        size_t mx = 2;
        for (auto& st : sites) mx = std::max(mx, (size_t)std::max(st.l, st.r) + 2);
        const size_t part = (mx / wqc::kVecMatChunk + 2) * mx;
        env_buf.ensure((total + part + count * 2) * sizeof(double2) + 4096);
        note_peak();
        double2* E = reinterpret_cast<double2*>(env_buf.p);
        vm_part = E + total;
        double2* amps = vm_part + part;
        HIPC(hipMemcpyAsync(E + loff[0], &kOne, sizeof(double2), hipMemcpyHostToDevice, stream));
        HIPC(hipMemcpyAsync(E + roff[n], &kOne, sizeof(double2), hipMemcpyHostToDevice, stream));
        std::vector<VecMatJob> jobs;
        for (int q = 0; q < n; ++q)
            jobs.push_back(VecMatJob{sites[q].ptr, E + loff[q], E + loff[q + 1], sites[q].l, sites[q].r, llen[q], 2, 0, 0,
                                     vm_part});
        for (int q = n - 1; q >= 0; --q)
            jobs.push_back(VecMatJob{sites[q].ptr, E + roff[q + 1], E + roff[q], sites[q].l, sites[q].r, rlen[q + 1], 2, 1,
                                     0, vm_part});
        VecMatJob* dj = upload(jobs);
        for (int q = 0; q < n; ++q)
            wqc::launch_vecmat(dj + q, sites[q].r, std::min(sites[q].l, llen[q]), 0, stream);
        for (int k = 0; k < n; ++k) {
            const int q = n - 1 - k;
            wqc::launch_vecmat(dj + n + k, sites[q].l, std::min(sites[q].r, rlen[q + 1]), 1, stream);
        }
        klaunch[5] += 2 * n + 1;
        std::vector<wqc::SiteAmpJob> sj(count);
        for (size_t k = 0; k < count; ++k) {
            const int t = targets[k];
            sj[k] = wqc::SiteAmpJob{E + loff[t], sites[t].ptr, E + roff[t + 1], amps + 2 * k, llen[t], sites[t].l,
                                    sites[t].r, rlen[t + 1]};
        }
        wqc::launch_site_amps(upload(sj), (int)count, stream);
        check_launch();
        std::vector<double2> h(2 * count);
        HIPC(hipMemcpyAsync(h.data(), amps, 2 * count * sizeof(double2), hipMemcpyDeviceToHost, stream));
        sync();
        for (size_t k = 0; k < count; ++k) {
            out[4 * k] = h[2 * k].x;
            out[4 * k + 1] = h[2 * k].y;
            out[4 * k + 2] = h[2 * k + 1].x;
            out[4 * k + 3] = h[2 * k + 1].y;
        }
    }
    DevBuf env_buf;

    // Left norm environments L_q (sites < q) and conjugated right norm environments Rb_q
    // (sites >= q), q = 0..n, each l_q x l_q at env_off[q]; L may be null. FL and HR are
    // scratch of 2 maxb^2 elements. L_{q+1} = A_stack^H (L_q A_flat) and
    // Rb_q = (A_stack Rb_{q+1})_flat A_flat^H, where A_flat is the site as (l) x (2r) and
    // A_stack the same memory as (2l) x (r).
    void build_envs(double2* L, double2* Rb, const std::vector<size_t>& env_off, double2* FL, double2* HR) {
        const double2 one = make_double2(1, 0), zero = make_double2(0, 0);
        if (L) HIPC(hipMemcpyAsync(L + env_off[0], &kOne, sizeof(double2), hipMemcpyHostToDevice, stream));
        HIPC(hipMemcpyAsync(Rb + env_off[n], &kOne, sizeof(double2), hipMemcpyHostToDevice, stream));
        for (int i = 0; i < n; ++i) {
            tables_reset();
            const int ql = i, qr = n - 1 - i;
            const int l1 = sites[ql].l, r1 = sites[ql].r, l2 = sites[qr].l, r2 = sites[qr].r;
            std::vector<GemmJob> a, b;
            if (L) {
                a.push_back(GemmJob{L + env_off[ql], sites[ql].ptr, FL, l1, 2 * r1, l1, l1, 2 * r1, 2 * r1, 0, 0, one, zero});
                b.push_back(GemmJob{sites[ql].ptr, FL, L + env_off[ql + 1], r1, r1, 2 * l1, r1, r1, r1, 1, 0, one, zero});
            }
            a.push_back(GemmJob{sites[qr].ptr, Rb + env_off[qr + 1], HR, 2 * l2, r2, r2, r2, r2, r2, 0, 0, one, zero});
            b.push_back(GemmJob{HR, sites[qr].ptr, Rb + env_off[qr], l2, l2, 2 * r2, 2 * r2, 2 * r2, l2, 0, 1, one, zero});
            wqc::launch_gemm(upload(a), (int)a.size(), std::max(L ? l1 : 1, 2 * l2), std::max(L ? 2 * r1 : 1, r2), stream);
            wqc::launch_gemm(upload(b), (int)b.size(), std::max(L ? r1 : 1, l2), std::max(L ? r1 : 1, l2), stream);
            check_launch();
            sync();
        }
    }

    // Terminal sampling (sample_terminal_measurements, src/sample.rs:192-211).
    //  exact: the reference's semantics, n <= 20: device projection to the statevector, then
    //         the snapped Born table and the seeded ChaCha12 draw on the host (born.cpp).
    //  MPS:   any n. Shot k draws qubits 0..n-1 in order from the MPS's conditional Born
    //         probabilities p(s | earlier outcomes) = <v A_s| R_{q+1} |v A_s>, with the right
    //         norm environments R and the carried left vector v renormalized by sqrt(p_s);
    //         the uniform for (shot k, qubit q) is draw k*n + q of the same seeded stream.
    //         Shots are exact samples of |psi_MPS|^2 / <psi|psi>.
    // The checks of compute_outcome_probabilities (src/sample.rs:91-109).
    void check_measures(const int32_t* mq, const int32_t* mc, size_t nm, size_t cbits) const {
        if (cbits == 0) fail(WQC_TN_EXEC_FAILED, "classical_bit_count must be > 0 for sample_counts");
        if (cbits > 64) fail(WQC_TN_INVALID_ARG, "classical_bit_count above 64 is not supported");
        for (size_t k = 0; k < nm; ++k) {
            if (mq[k] < 0 || mq[k] >= n)
                fail(WQC_TN_INVALID_ARG, "qubit index " + std::to_string(mq[k]) + " out of bounds (limit: " +
                                             std::to_string(n) + ")");
            if (mc[k] < 0 || (size_t)mc[k] >= cbits)
                fail(WQC_TN_EXEC_FAILED, "classical bit index " + std::to_string(mc[k]) +
                                             " out of bounds (limit: " + std::to_string(cbits) + ")");
        }
    }

    // compute_outcome_probabilities (src/sample.rs:85-131) of this state: the statevector
    // projected on the device, the snapped Born table on the host (born.cpp), key order.
    std::vector<wqc::BornEntry> outcome_probabilities(const int32_t* mq, const int32_t* mc, size_t nm,
                                                      size_t cbits) {
        check_measures(mq, mc, nm, cbits);
        if (n > 20) fail(WQC_TN_EXEC_FAILED, "full statevector projection requires qubit_count <= 20");
        std::vector<double> sv((size_t)2 << n);
        statevector(sv.data(), (size_t)1 << n);
        return wqc::born_table(sv.data(), n, mq, mc, nm, cbits);
    }

    std::vector<std::pair<uint64_t, uint64_t>> sample_counts(const int32_t* mq, const int32_t* mc,
                                                             size_t nm, size_t cbits, uint64_t shots,
                                                             uint64_t seed, int method) {
        check_measures(mq, mc, nm, cbits);
        const bool exact = method == 0 || (method == 2 && n <= 20);
        if (method < 0 || method > 2) fail(WQC_TN_INVALID_ARG, "unknown sampling method");
        if (exact) {
            return wqc::draw_counts(outcome_probabilities(mq, mc, nm, cbits), shots, seed);
        }
        if (n > 64) fail(WQC_TN_INVALID_ARG, "the MPS sampler supports qubit_count <= 64");
        flush();
        sync();
        size_t maxb = 1, env_elems = 0;
        std::vector<size_t> env_off(n + 1);
        for (int q = 0; q <= n; ++q) {
            const size_t l = q < n ? sites[q].l : 1;
            env_off[q] = env_elems;
            env_elems += l * l;
            maxb = std::max(maxb, l);
        }
        const uint64_t chunk = std::min<uint64_t>(std::max<uint64_t>(shots, 1), 16384);
        const size_t fl = 2 * maxb * maxb;
        // Rb, FL, HR, then per chunk: V, Vn (chunk x maxb), W (chunk x 2 maxb), Z (2 chunk x maxb),
        // uniforms (chunk x n doubles) and bits.
        const size_t vec = (size_t)chunk * maxb;
        const size_t total = env_elems + 2 * fl + 2 * vec + 2 * vec + 2 * vec + ((size_t)chunk * n + 1) / 2 + chunk + 64;
        ews.ensure(total * sizeof(double2));
        note_peak();
        double2* Rb = (double2*)ews.p;
        double2* FL = Rb + env_elems;
        double2* HR = FL + fl;
        double2* V = HR + fl;
        double2* Vn = V + vec;
        double2* W = Vn + vec;
        double2* Z = W + 2 * vec;
        double* U = reinterpret_cast<double*>(Z + 2 * vec);
        unsigned long long* B = reinterpret_cast<unsigned long long*>(U + (size_t)chunk * n + 2);
        build_envs(nullptr, Rb, env_off, FL, HR);
        const double2 one = make_double2(1, 0), zero = make_double2(0, 0);
        wqc::ChaCha12Stream rng(seed);
        std::map<uint64_t, uint64_t> counts;
        std::vector<double> hu;
        std::vector<unsigned long long> hb;
        std::vector<double2> ones;
        for (uint64_t s0 = 0; s0 < shots; s0 += chunk) {
            const long m = (long)std::min<uint64_t>(chunk, shots - s0);
            hu.resize((size_t)m * n);
            for (auto& x : hu) x = rng.next_f64();
            HIPC(hipMemcpyAsync(U, hu.data(), hu.size() * sizeof(double), hipMemcpyHostToDevice, stream));
            wqc::launch_fill_int(reinterpret_cast<int*>(B), 0, (int)(2 * m), stream);
            ones.assign((size_t)m, make_double2(1, 0));
            HIPC(hipMemcpyAsync(V, ones.data(), (size_t)m * sizeof(double2), hipMemcpyHostToDevice, stream));
            for (int q = 0; q < n; ++q) {
                check_cancel();
                tables_reset();
                const int l = sites[q].l, r = sites[q].r;
                std::vector<GemmJob> a{GemmJob{V, sites[q].ptr, W, (int)m, 2 * r, l, l, 2 * r, 2 * r, 0, 0, one, zero}};
                std::vector<GemmJob> b{GemmJob{W, Rb + env_off[q + 1], Z, (int)m, r, r, 2 * r, r, r, 0, 1, one, zero},
                                       GemmJob{W + r, Rb + env_off[q + 1], Z + (size_t)m * r, (int)m, r, r, 2 * r, r, r, 0, 1, one, zero}};
                wqc::launch_gemm(upload(a), 1, (int)m, 2 * r, stream);
                wqc::launch_gemm(upload(b), 2, (int)m, r, stream);
                wqc::launch_sample_step(W, Z, r, m, U + q, n, q, B, Vn, stream);
                check_launch();
                std::swap(V, Vn);
                sync();
            }
            hb.resize((size_t)m);
            HIPC(hipMemcpyAsync(hb.data(), B, (size_t)m * sizeof(unsigned long long), hipMemcpyDeviceToHost, stream));
            sync();
            for (long k = 0; k < m; ++k) {
                uint64_t key = 0;
                for (size_t j = 0; j < nm; ++j) {
                    const uint64_t bit = (hb[k] >> mq[j]) & 1u;
                    const uint64_t msk = (uint64_t)1 << mc[j];
                    key = bit ? (key | msk) : (key & ~msk);
                }
                ++counts[key];
            }
        }
        return std::vector<std::pair<uint64_t, uint64_t>>(counts.begin(), counts.end());
    }

    // ------------------------------------------------------------ snapshots (gradient checkpoints)
    struct Snapshot {
        std::vector<double2*> ptr;
        std::vector<int> l, r;
        size_t bytes = 0;
        void release() {
            for (auto p : ptr)
                if (p) (void)hipFree(p);
            ptr.clear();
            bytes = 0;
        }
    };
    void snapshot(Snapshot& sn) {
        flush();
        sync();
        sn.release();
        sn.ptr.assign(n, nullptr);
        sn.l.assign(n, 1);
        sn.r.assign(n, 1);
        for (int i = 0; i < n; ++i) {
            const size_t cnt = (size_t)sites[i].l * 2 * sites[i].r;
            HIPC(hipMalloc(&sn.ptr[i], cnt * sizeof(double2)));
            HIPC(hipMemcpyAsync(sn.ptr[i], sites[i].ptr, cnt * sizeof(double2), hipMemcpyDeviceToDevice, stream));
            sn.l[i] = sites[i].l;
            sn.r[i] = sites[i].r;
            sn.bytes += cnt * sizeof(double2);
        }
        sync();
    }
    void restore(const Snapshot& sn) {
        flush();
        sync();
        for (int i = 0; i < n; ++i) {
            const size_t cnt = (size_t)sn.l[i] * 2 * sn.r[i];
            ensure_site_cap(i, cnt);
            HIPC(hipMemcpyAsync(sites[i].ptr, sn.ptr[i], cnt * sizeof(double2), hipMemcpyDeviceToDevice, stream));
            sites[i].l = sn.l[i];
            sites[i].r = sn.r[i];
        }
        sync();
    }

    // ------------------------------------------------------------ MPO x MPS
    // Sites of this state <- MPO x (sites of src). mpo[q] is D_q x D_{q+1} x 2 x 2 (row-major
    // complex); the left boundary keeps MPO row 0 ("initial") and the right boundary column 1
    // ("final").
    void set_mpo_product(const wqc_tn_state& src, const std::vector<std::vector<double2>>& mpo,
                         const std::vector<int>& D) {
        flush();
        sync();
        size_t wmax = 1;
        for (auto& w : mpo) wmax = std::max(wmax, w.size());
        DevBuf wbuf;
        wbuf.ensure(wmax * sizeof(double2));
        for (int q = 0; q < n; ++q) {
            const Site& a = src.sites[q];
            const int w0 = q == 0 ? 0 : 0, wn = q == 0 ? 1 : D[q];
            const int v0 = q == n - 1 ? 1 : 0, vn = q == n - 1 ? 1 : D[q + 1];
            const size_t cnt = (size_t)a.l * wn * 2 * a.r * vn;
            ensure_site_cap(q, cnt);
            HIPC(hipMemcpyAsync(wbuf.p, mpo[q].data(), mpo[q].size() * sizeof(double2), hipMemcpyHostToDevice, stream));
            wqc::launch_mpo_site(a.ptr, (const double2*)wbuf.p, sites[q].ptr, a.l, a.r, D[q], D[q + 1], w0, wn, v0, vn,
                                 stream);
            check_launch();
            sites[q].l = a.l * wn;
            sites[q].r = a.r * vn;
            sync();
        }
        wbuf.release();
    }

    // Site cotangents of Re <bra|ket> (this state the ket): cot[q] = scale * L_q^H (Bra_q Rb_{q+1}),
    // so that d Re <bra|ket> = Re sum conj(cot[q] / scale) . dKet_q over every site, with the
    // two-state left environments L_{q+1} = sum_s Bra_s^H L_q Ket_s (L_0 = 1) and the conjugated
    // right ones Rb_q = sum_s Bra_s Rb_{q+1} Ket_s^H (Rb_n = 1). cot[q] is sized to the ket's
    // site q.
    void site_cotangents(const wqc_tn_state& bra, std::vector<std::unique_ptr<DevBuf>>& cot, double scale) {
        flush();
        sync();
        const wqc_tn_state& ket = *this;
        size_t env = 0, maxbk = 1;
        std::vector<size_t> off(n + 1);
        auto lb = [&](int q) { return q < n ? bra.sites[q].l : 1; };
        auto lk = [&](int q) { return q < n ? ket.sites[q].l : 1; };
        for (int q = 0; q <= n; ++q) {
            off[q] = env;
            env += (size_t)lb(q) * lk(q);
            maxbk = std::max(maxbk, (size_t)std::max(lb(q), lk(q)));
        }
        const size_t fl = 2 * maxbk * maxbk;
        ews.ensure((2 * env + 2 * fl + 64) * sizeof(double2));
        note_peak();
        double2* L = (double2*)ews.p;
        double2* Rb = L + env;
        double2* FL = Rb + env;
        double2* HR = FL + fl;
        const double2 one = make_double2(1, 0), zero = make_double2(0, 0);
        HIPC(hipMemcpyAsync(L + off[0], &kOne, sizeof(double2), hipMemcpyHostToDevice, stream));
        HIPC(hipMemcpyAsync(Rb + off[n], &kOne, sizeof(double2), hipMemcpyHostToDevice, stream));
        for (int i = 0; i < n; ++i) {
            tables_reset();
            const int ql = i, qr = n - 1 - i;
            const Site &bl = bra.sites[ql], &kl = ket.sites[ql], &br = bra.sites[qr], &kr = ket.sites[qr];
            std::vector<GemmJob> a{
                GemmJob{L + off[ql], kl.ptr, FL, bl.l, 2 * kl.r, kl.l, kl.l, 2 * kl.r, 2 * kl.r, 0, 0, one, zero},
                GemmJob{br.ptr, Rb + off[qr + 1], HR, 2 * br.l, kr.r, br.r, br.r, kr.r, kr.r, 0, 0, one, zero}};
            std::vector<GemmJob> b{
                GemmJob{bl.ptr, FL, L + off[ql + 1], bl.r, kl.r, 2 * bl.l, bl.r, kl.r, kl.r, 1, 0, one, zero},
                GemmJob{HR, kr.ptr, Rb + off[qr], br.l, kr.l, 2 * kr.r, 2 * kr.r, 2 * kr.r, kr.l, 0, 1, one, zero}};
            wqc::launch_gemm(upload(a), 2, std::max(bl.l, 2 * br.l), std::max(2 * kl.r, kr.r), stream);
            wqc::launch_gemm(upload(b), 2, std::max(bl.r, br.l), std::max(kl.r, kr.l), stream);
            check_launch();
            sync();
        }
        cot.resize(n);
        for (int q = 0; q < n; ++q) {
            tables_reset();
            const Site &bs = bra.sites[q], &ks = ket.sites[q];
            if (!cot[q]) cot[q] = std::make_unique<DevBuf>();
            cot[q]->ensure((size_t)ks.l * 2 * ks.r * sizeof(double2));
            double2* c = (double2*)cot[q]->p;
            // G = Bra_q Rb_{q+1}: (bra.l, 2, ket.r); cot = scale L_q^H G: (ket.l, 2, ket.r).
            std::vector<GemmJob> g1{gj(bs.ptr, bs.r, 0, Rb + off[q + 1], ks.r, 0, FL, ks.r, 2 * bs.l, ks.r, bs.r, 1.0, 0.0)};
            std::vector<GemmJob> g2{gj(L + off[q], ks.l, 1, FL, 2 * ks.r, 0, c, 2 * ks.r, ks.l, 2 * ks.r, bs.l, scale, 0.0)};
            gemm(g1, 5);
            gemm(g2, 5);
            sync();
        }
    }

    // <psi|P_k|psi> of Pauli strings by transfer matrices. The reference evaluates the same
    // sum over the dense statevector, sum_i conj(psi_P(i)) phase_i psi_i
    // (src/expectation.rs:140-167, n <= 20 through contract_to_statevector); here any n.
    // Site q carries qubit q, whose Pauli is the label's character n-1-q (Qiskit order,
    // src/expectation.rs:98-132). Left norm environments L_q (sites < q) and conjugated right
    // environments Rb_q (sites >= q) are built once; a term runs the transfer
    // E <- B_stack^H (E A_flat), B_t = sum_s' conj(P[s',t]) A_s', only over the sites from its
    // first to its last non-identity character and closes with sum E conj(Rb).
    void pauli_expectations(const char* labels, size_t nterms, double* out_re, double* out_im) {
        flush();
        sync();
        std::vector<int> lo(nterms, -1), hi(nterms, -1);
        for (size_t k = 0; k < nterms; ++k) {
            const char* lab = labels + k * (size_t)n;
            for (int c = 0; c < n; ++c) {
                const char ch = lab[c];
                if (ch != 'I' && ch != 'X' && ch != 'Y' && ch != 'Z')
                    fail(WQC_TN_INVALID_ARG, std::string("invalid pauli character ") +
                                                 (ch ? std::string("'") + ch + "'" : std::string("NUL")) +
                                                 " in label " + std::to_string(k));
            }
            for (int q = 0; q < n; ++q)
                if (lab[n - 1 - q] != 'I') {
                    if (lo[k] < 0) lo[k] = q;
                    hi[k] = q;
                }
        }
        auto ldim = [&](int q) { return q < n ? sites[q].l : 1; };
        size_t maxb = 1, max_site = 1, env_elems = 0;
        std::vector<size_t> env_off(n + 1);
        for (int q = 0; q <= n; ++q) {
            env_off[q] = env_elems;
            env_elems += (size_t)ldim(q) * ldim(q);
            maxb = std::max(maxb, (size_t)ldim(q));
        }
        for (int q = 0; q < n; ++q) max_site = std::max(max_site, (size_t)sites[q].l * 2 * sites[q].r);
        const size_t sq = maxb * maxb, fl = 2 * maxb * maxb;
        // Terms per pass bound the per-term buffers (two environments and one F) to 2 GiB.
        const size_t per_term = 2 * sq + fl;
        const size_t chunk = std::max<size_t>(1, std::min<size_t>(nterms ? nterms : 1, ((size_t)2 << 30) / (per_term * sizeof(double2))));
        const size_t fixed = 2 * env_elems + 2 * fl + 3 * max_site + 2 * nterms + 64;
        ews.ensure((fixed + chunk * per_term) * sizeof(double2));
        note_peak();
        double2* base = (double2*)ews.p;
        double2* L = base;
        double2* Rb = L + env_elems;
        double2* FL = Rb + env_elems;
        double2* HR = FL + fl;
        double2* BX = HR + fl;
        double2* BY = BX + max_site;
        double2* BZ = BY + max_site;
        double2* res = BZ + max_site;
        double2* terms = res + 2 * nterms + 64;
        auto site_flat = [&](int q) { return sites[q].ptr; };
        const double2 one = make_double2(1, 0), zero = make_double2(0, 0);
        build_envs(L, Rb, env_off, FL, HR);
        for (size_t k0 = 0; k0 < nterms; k0 += chunk) {
            const size_t k1 = std::min(nterms, k0 + chunk);
            std::vector<double2*> cur(k1 - k0, nullptr);
            for (int q = 0; q < n; ++q) {
                std::vector<size_t> act;
                for (size_t k = k0; k < k1; ++k)
                    if (lo[k] >= 0 && lo[k] <= q && q <= hi[k]) act.push_back(k);
                if (act.empty()) continue;
                tables_reset();
                const int l = sites[q].l, r = sites[q].r;
                wqc::launch_pauli_site(site_flat(q), BX, BY, BZ, l, r, stream);
                std::vector<GemmJob> a, b;
                for (size_t k : act) {
                    double2* tb = terms + (k - k0) * per_term;
                    double2* e0 = tb;
                    double2* e1 = tb + sq;
                    double2* f = tb + 2 * sq;
                    const double2* ein = q == lo[k] ? L + env_off[q] : cur[k - k0];
                    double2* eout = (ein == e0) ? e1 : e0;
                    const char ch = labels[k * (size_t)n + (n - 1 - q)];
                    const double2* bsite = ch == 'X' ? BX : ch == 'Y' ? BY : ch == 'Z' ? BZ : site_flat(q);
                    a.push_back(GemmJob{ein, site_flat(q), f, l, 2 * r, l, l, 2 * r, 2 * r, 0, 0, one, zero});
                    b.push_back(GemmJob{bsite, f, eout, r, r, 2 * l, r, r, r, 1, 0, one, zero});
                    cur[k - k0] = eout;
                }
                wqc::launch_gemm(upload(a), (int)a.size(), l, 2 * r, stream);
                wqc::launch_gemm(upload(b), (int)b.size(), r, r, stream);
                check_launch();
                sync();
            }
            tables_reset();
            std::vector<wqc::CloseJob> cj;
            for (size_t k = k0; k < k1; ++k) {
                if (lo[k] < 0)
                    cj.push_back(wqc::CloseJob{L + env_off[n], Rb + env_off[n], res + k, 1});
                else {
                    const int qe = hi[k] + 1;
                    cj.push_back(wqc::CloseJob{cur[k - k0], Rb + env_off[qe], res + k, (long)ldim(qe) * ldim(qe)});
                }
            }
            wqc::launch_env_close(upload(cj), (int)cj.size(), stream);
            check_launch();
            sync();
        }
        std::vector<double2> h(std::max<size_t>(nterms, 1));
        if (nterms) HIPC(hipMemcpyAsync(h.data(), res, nterms * sizeof(double2), hipMemcpyDeviceToHost, stream));
        sync();
        for (size_t k = 0; k < nterms; ++k) {
            out_re[k] = h[k].x;
            if (out_im) out_im[k] = h[k].y;
        }
    }

    // contract_to_statevector (src/tn/mps.rs:215-260).
    void statevector(double* out, size_t len_complex) {
        flush();
        sync();
        if (n > 20) fail(WQC_TN_EXEC_FAILED, "full statevector projection requires qubit_count <= 20");
        const size_t dim = (size_t)1 << n;
        if (len_complex < dim) fail(WQC_TN_INVALID_ARG, "statevector buffer too small");
        size_t mx = 2;
        for (int i = 0; i < n; ++i) mx = std::max(mx, (size_t)sites[i].r * ((size_t)2 << i));
        for (int i = 0; i < n; ++i) mx = std::max(mx, (size_t)sites[i].l * ((size_t)1 << i));
        ws.ensure(2 * mx * sizeof(double2));
        note_peak();
        double2* a = (double2*)ws.p;
        double2* b = a + mx;
        HIPC(hipMemcpyAsync(a, &kOne, sizeof(double2), hipMemcpyHostToDevice, stream));
        for (int i = 0; i < n; ++i) {
            const long prefix = 1L << i;
            if (i > 0 && sites[i - 1].r != sites[i].l)
                fail(WQC_TN_EXEC_FAILED, "MPS bond layout mismatch during statevector projection");
            if (i == 0 && sites[0].l != 1)
                fail(WQC_TN_EXEC_FAILED, "MPS bond layout mismatch during statevector projection");
            wqc::launch_statevector_step(a, sites[i].ptr, b, sites[i].l, sites[i].r, prefix, stream);
            check_launch();
            std::swap(a, b);
        }
        if (sites[n - 1].r != 1) fail(WQC_TN_EXEC_FAILED, "projected statevector length mismatch");
        HIPC(hipMemcpyAsync(out, a, dim * sizeof(double2), hipMemcpyDeviceToHost, stream));
        sync();
    }
};

constexpr double2 wqc_tn_state::kOne;

// =====================================================================================
// C ABI
// =====================================================================================
extern "C" {

const char* wqc_tn_last_error(void) { return g_last_error.c_str(); }

const char* wqc_tn_backend_from_env(void) {
    const char* v = std::getenv("WQC_TN_BACKEND");
    if (!v) return "cpu";
    std::string s(v);
    for (auto& ch : s) ch = (char)std::tolower((unsigned char)ch);
    if (s == "hip" || s == "gpu" || s == "rocm") return "hip";
    return "cpu";
}

int wqc_tn_engine_status(char* active, size_t cap, uint64_t* vram_budget_bytes) {
    return guarded([&] {
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || count == 0) fail(WQC_TN_NO_DEVICE, "no HIP device");
        hipDeviceProp_t prop;
        HIPC(hipGetDeviceProperties(&prop, 0));
        std::string s = std::string("hip ") + prop.gcnArchName;
        if (active && cap) {
            std::strncpy(active, s.c_str(), cap - 1);
            active[cap - 1] = 0;
        }
        size_t fr = 0, tot = 0;
        HIPC(hipSetDevice(0));
        HIPC(hipMemGetInfo(&fr, &tot));
        if (vram_budget_bytes) *vram_budget_bytes = fr;
    });
}

int wqc_tn_device_open(int ordinal, wqc_tn_device** out) {
    return guarded([&] {
        if (!out) fail(WQC_TN_INVALID_ARG, "null output");
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || count == 0) fail(WQC_TN_NO_DEVICE, "no HIP device");
        if (ordinal < 0 || ordinal >= count) fail(WQC_TN_INVALID_ARG, "device ordinal out of range");
        auto d = std::make_unique<wqc_tn_device>();
        d->ordinal = ordinal;
        HIPC(hipSetDevice(ordinal));
        HIPC(hipGetDeviceProperties(&d->prop, ordinal));
        if (std::strncmp(d->prop.gcnArchName, "gfx950", 6) != 0)
            fail(WQC_TN_NO_DEVICE, std::string("device is ") + d->prop.gcnArchName + ", this build targets gfx950");
        // The scratch array's address on this device (each device has its own copy).
        void* sp = nullptr;
        HIPC(hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch)));
        if (!sp) fail(WQC_TN_EXEC_FAILED, "scratch array has no device address");
        d->scratch = static_cast<double2*>(sp);
        if (const char* e = std::getenv("WQC_TN_TUNE_CACHE")) d->tune_path = e;
        d->load_tune_cache();
        *out = d.release();
    });
}

void wqc_tn_device_close(wqc_tn_device* dev) { delete dev; }

int wqc_tn_state_new(wqc_tn_device* dev, size_t qubit_count, size_t max_bond_dim, int precision,
                     wqc_tn_state** out) {
    return guarded([&] {
        if (!dev || !out) fail(WQC_TN_INVALID_ARG, "null argument");
        if (qubit_count == 0) fail(WQC_TN_INVALID_ARG, "invalid qubit count: 0");
        if (max_bond_dim == 0) fail(WQC_TN_EXEC_FAILED, "MPS max bond dimension must be > 0");
        if (qubit_count > (1u << 20)) fail(WQC_TN_INVALID_ARG, "qubit count too large");
        if (!valid_precision(precision)) fail(WQC_TN_INVALID_ARG, "unknown precision mode");
        // ensure_mps_workspace_budget analogue (src/tn/mps.rs:26-48): N chi^2 32 bytes
        // against free device memory (the blocks the device's cache holds count as free).
        HIPC(hipSetDevice(dev->ordinal));
        size_t fr = 0, tot = 0;
        HIPC(hipMemGetInfo(&fr, &tot));
        fr += dev->pool.held_bytes();
        unsigned __int128 need = (unsigned __int128)qubit_count * max_bond_dim * max_bond_dim * 32;
        if (need > fr)
            fail(WQC_TN_OUT_OF_MEMORY, "insufficient device memory: required " +
                                           std::to_string((unsigned long long)std::min<unsigned __int128>(need, ~0ull)) +
                                           " bytes, available " + std::to_string(fr));
        auto st = std::make_unique<wqc_tn_state>();
        st->init(dev, qubit_count, max_bond_dim, precision);
        *out = st.release();
    });
}

void wqc_tn_state_free(wqc_tn_state* st) { delete st; }

size_t wqc_tn_qubit_count(const wqc_tn_state* st) { return st ? (size_t)st->n : 0; }

int wqc_tn_set_max_bond_dim(wqc_tn_state* st, size_t max_bond_dim) {
    return guarded([&] {
        if (!st || max_bond_dim == 0) fail(WQC_TN_INVALID_ARG, "bad max bond dim");
        st->flush();
        st->chi = max_bond_dim;
    });
}

int wqc_tn_set_reject_truncation(wqc_tn_state* st, int reject) {
    return guarded([&] {
        if (!st) fail(WQC_TN_INVALID_ARG, "null state");
        st->flush();
        st->reject = reject != 0;
    });
}

int wqc_tn_apply_gate(wqc_tn_state* st, const wqc_tn_gate* gate) {
    return guarded([&] {
        if (!st || !gate) fail(WQC_TN_INVALID_ARG, "null argument");
        HIPC(hipSetDevice(st->dev->ordinal));
        st->queue_gate(*gate);
        st->flush();
    });
}

int wqc_tn_apply_circuit(wqc_tn_state* st, const wqc_tn_gate* gates, size_t count) {
    return guarded([&] {
        if (!st || (!gates && count)) fail(WQC_TN_INVALID_ARG, "null argument");
        HIPC(hipSetDevice(st->dev->ordinal));
        try {
            st->queue_circuit(gates, count);
            st->flush();
        } catch (...) {
            st->pend1.clear();
            st->pend2.clear();
            std::fill(st->touched.begin(), st->touched.end(), 0);
            throw;
        }
    });
}

int wqc_tn_overlap(wqc_tn_state* ket, wqc_tn_state* bra, double out[2]) {
    return guarded([&] {
        if (!ket || !bra || !out) fail(WQC_TN_INVALID_ARG, "null argument");
        if (ket->dev != bra->dev) fail(WQC_TN_INVALID_ARG, "states on different devices");
        HIPC(hipSetDevice(ket->dev->ordinal));
        ket->overlap_with(*bra, out);
    });
}

namespace {
// adjoint_gate_sequence (src/tn/mps.rs:1398-1419): each gate's adjoint as gates of the same
// set; S^H = Z S and T^H = Z T T T are applied as S then Z and T, T, T then Z.
void append_adjoint(const wqc_tn_gate& g, std::vector<wqc_tn_gate>& out) {
    wqc_tn_gate a = g;
    switch (g.kind) {
        case WQC_TN_GATE_RX:
        case WQC_TN_GATE_RY:
        case WQC_TN_GATE_RZ:
            a.angle = -g.angle;
            out.push_back(a);
            break;
        case WQC_TN_GATE_S:
        case WQC_TN_GATE_T: {
            const int reps = g.kind == WQC_TN_GATE_S ? 1 : 3;
            for (int r = 0; r < reps; ++r) out.push_back(g);
            wqc_tn_gate z = g;
            z.kind = WQC_TN_GATE_Z;
            out.push_back(z);
            break;
        }
        case WQC_TN_GATE_MEASURE:
        case WQC_TN_GATE_RESET:
        case WQC_TN_GATE_IF:
            fail(WQC_TN_EXEC_FAILED, "time-split overlap does not apply measurement or reset");
        default:
            out.push_back(g);  // H, X, Y, Z, CNOT, CZ, CCNOT are their own adjoints
    }
}
}  // namespace

int wqc_tn_contract_time_split(wqc_tn_device* dev, size_t qubit_count, const wqc_tn_gate* gates, size_t count,
                               size_t split, size_t max_bond_dim, int precision, double out[2]) {
    return guarded([&] {
        if (!dev || !out || (!gates && count)) fail(WQC_TN_INVALID_ARG, "null argument");
        if (split > count) fail(WQC_TN_EXEC_FAILED, "time-split index past the gate list");
        auto make = [&]() {
            wqc_tn_state* p = nullptr;
            const int rc = wqc_tn_state_new(dev, qubit_count, max_bond_dim, precision, &p);
            if (rc != WQC_TN_OK) fail(rc, g_last_error);
            p->reject = true;
            return std::unique_ptr<wqc_tn_state, void (*)(wqc_tn_state*)>(p, wqc_tn_state_free);
        };
        HIPC(hipSetDevice(dev->ordinal));
        auto ket = make();
        for (size_t i = 0; i < split; ++i) ket->queue_gate(gates[i]);
        ket->flush();
        std::vector<wqc_tn_gate> bra_gates;
        for (size_t i = count; i-- > split;) append_adjoint(gates[i], bra_gates);
        auto bra = make();
        for (const auto& g : bra_gates) bra->queue_gate(g);
        bra->flush();
        ket->overlap_with(*bra, out);
    });
}

int wqc_tn_apply_adjacent_disjoint_gates(wqc_tn_state* st, const wqc_tn_gate* gates, size_t count) {
    return guarded([&] {
        if (!st || (!gates && count)) fail(WQC_TN_INVALID_ARG, "null argument");
        HIPC(hipSetDevice(st->dev->ordinal));
        st->check_cancel();
        if (count < 2) {
            for (size_t i = 0; i < count; ++i) st->queue_gate(gates[i]);
            st->flush();
            return;
        }
        std::vector<char> seen(st->n, 0);
        for (size_t i = 0; i < count; ++i) {
            const wqc_tn_gate& g = gates[i];
            if (g.kind != WQC_TN_GATE_CNOT && g.kind != WQC_TN_GATE_CZ)
                fail(WQC_TN_EXEC_FAILED, "disjoint layer contains a gate that is not CNOT or CZ");
            st->check_gate(g);
            const int p = std::min(g.q0, g.q1), q = std::max(g.q0, g.q1);
            if (q != p + 1) fail(WQC_TN_INVALID_ARG, "disjoint layer gate is not adjacent");
            if (seen[p] || seen[q]) fail(WQC_TN_INVALID_ARG, "disjoint layer gates share a qubit");
            seen[p] = seen[q] = 1;
        }
        st->flush();
        for (size_t i = 0; i < count; ++i) {
            Mat4 u;
            two_qubit_matrix(gates[i].kind, u);
            st->queue_unit(std::min(gates[i].q0, gates[i].q1), u, gates[i].q0 < gates[i].q1);
        }
        st->flush();
    });
}

int wqc_tn_apply_unary_at(wqc_tn_state* st, size_t site, const wqc_tn_gate* gate) {
    return guarded([&] {
        if (!st || !gate || site >= (size_t)st->n || !is_one_qubit(gate->kind))
            fail(WQC_TN_INVALID_ARG, "bad unary_at arguments");
        Mat2 u;
        unary_matrix(gate->kind, gate->angle, u);
        st->queue_one((int)site, u);
        st->flush();
    });
}

int wqc_tn_apply_tdg_at(wqc_tn_state* st, size_t site) {
    return guarded([&] {
        if (!st || site >= (size_t)st->n) fail(WQC_TN_INVALID_ARG, "bad site");
        Mat2 u;
        unary_matrix(0, 0, u);
        st->queue_one((int)site, u);
        st->flush();
    });
}

static int pair_at(wqc_tn_state* st, size_t site, int kind, int ctrl_on_left) {
    return guarded([&] {
        if (!st || site + 1 >= (size_t)st->n) fail(WQC_TN_INVALID_ARG, "bad site");
        Mat4 u;
        two_qubit_matrix(kind, u);
        st->queue_unit((int)site, u, ctrl_on_left != 0);
        st->flush();
    });
}

int wqc_tn_apply_swap_at(wqc_tn_state* st, size_t site) { return pair_at(st, site, KIND_SWAP, 1); }

int wqc_tn_apply_pair_at(wqc_tn_state* st, size_t site, const wqc_tn_gate* gate, int ctrl_on_left) {
    if (!gate || (gate->kind != WQC_TN_GATE_CNOT && gate->kind != WQC_TN_GATE_CZ)) {
        g_last_error = "apply_pair_at takes CNOT or CZ";
        return WQC_TN_INVALID_ARG;
    }
    return pair_at(st, site, gate->kind, ctrl_on_left);
}

int wqc_tn_apply_cnot_at(wqc_tn_state* st, size_t site, int ctrl_on_left) {
    return pair_at(st, site, WQC_TN_GATE_CNOT, ctrl_on_left);
}

void wqc_tn_request_cancel(wqc_tn_state* st) {
    if (st) st->cancel.store(1);
}

void wqc_tn_clear_cancel(wqc_tn_state* st) {
    if (st) st->cancel.store(0);
}

int wqc_tn_amplitude_at_compact_zero(wqc_tn_state* st, double out[2]) {
    return guarded([&] {
        if (!st || !out) fail(WQC_TN_INVALID_ARG, "null argument");
        auto a = st->amplitude_zero_timed();
        out[0] = a.first;
        out[1] = a.second;
    });
}

int wqc_tn_amplitude_at_compact_zero_noalloc(wqc_tn_state* st, double out[2]) {
    return guarded([&] {
        if (!st || !out) fail(WQC_TN_INVALID_ARG, "null argument");
        HIPC(hipSetDevice(st->dev->ordinal));
        auto a = st->amplitude_zero_noalloc();
        out[0] = a.first;
        out[1] = a.second;
    });
}

int wqc_tn_site_amplitudes(wqc_tn_state* st, size_t target, double out[4]) {
    return guarded([&] {
        if (!st || !out || target >= (size_t)st->n) fail(WQC_TN_INVALID_ARG, "bad target");
        st->site_amplitudes((int)target, out);
    });
}

int wqc_tn_site_amplitudes_batch(wqc_tn_state* st, const int32_t* targets, size_t count, double* out) {
    return guarded([&] {
        if (!st || (count && (!targets || !out))) fail(WQC_TN_INVALID_ARG, "null argument");
        for (size_t k = 0; k < count; ++k)
            if (targets[k] < 0 || targets[k] >= st->n) fail(WQC_TN_INVALID_ARG, "bad target");
        if (count) st->site_amplitudes_batch(targets, count, out);
    });
}

int wqc_tn_control_probability(wqc_tn_state* st, size_t control, double* out) {
    return guarded([&] {
        if (!st || !out || control >= (size_t)st->n) fail(WQC_TN_INVALID_ARG, "bad control");
        double a[4];
        st->site_amplitudes((int)control, a);
        *out = a[2] * a[2] + a[3] * a[3];
    });
}

int wqc_tn_pauli_expectations(wqc_tn_state* st, const char* labels, size_t n_terms, double* out,
                              double* out_imag) {
    return guarded([&] {
        if (!st || (n_terms && (!labels || !out))) fail(WQC_TN_INVALID_ARG, "null argument");
        st->pauli_expectations(labels, n_terms, out, out_imag);
    });
}

int wqc_tn_observable_expectation(wqc_tn_state* st, const char* labels, const double* coeffs,
                                  size_t n_terms, double out[2]) {
    return guarded([&] {
        if (!st || !out || (n_terms && (!labels || !coeffs))) fail(WQC_TN_INVALID_ARG, "null argument");
        if (!n_terms) fail(WQC_TN_INVALID_ARG, "observable must have at least one term");
        std::vector<double> v(n_terms);
        st->pauli_expectations(labels, n_terms, v.data(), nullptr);
        // observable_expectation (src/expectation.rs:170-186): sum of coeff * <P>, in term order;
        // compute_expectations snaps each component to the 1e-6 grid (src/engine.rs:30-40).
        double re = 0.0, im = 0.0;
        for (size_t k = 0; k < n_terms; ++k) {
            re += coeffs[2 * k] * v[k];
            im += coeffs[2 * k + 1] * v[k];
        }
        out[0] = std::round(re * 1e6) / 1e6;
        out[1] = std::round(im * 1e6) / 1e6;
    });
}

int wqc_tn_sample_counts(wqc_tn_state* st, const int32_t* measure_qubits, const int32_t* measure_cbits,
                         size_t n_measures, size_t classical_bits, uint64_t shots, uint64_t seed,
                         int method, uint64_t* keys, uint64_t* counts, size_t cap, size_t* n_out) {
    return guarded([&] {
        if (!st || !n_out || (n_measures && (!measure_qubits || !measure_cbits)) || (cap && (!keys || !counts)))
            fail(WQC_TN_INVALID_ARG, "null argument");
        auto res = st->sample_counts(measure_qubits, measure_cbits, n_measures, classical_bits, shots, seed, method);
        if (res.size() > cap) fail(WQC_TN_INVALID_ARG, "output capacity too small for the distinct outcomes");
        for (size_t i = 0; i < res.size(); ++i) {
            keys[i] = res[i].first;
            counts[i] = res[i].second;
        }
        *n_out = res.size();
    });
}

int wqc_tn_outcome_probabilities(wqc_tn_state* st, const int32_t* measure_qubits, const int32_t* measure_cbits,
                                 size_t n_measures, size_t classical_bits, uint64_t* keys, double* probs, size_t cap,
                                 size_t* n_out) {
    return guarded([&] {
        if (!st || !n_out || (n_measures && (!measure_qubits || !measure_cbits)) || (cap && (!keys || !probs)))
            fail(WQC_TN_INVALID_ARG, "null argument");
        auto res = st->outcome_probabilities(measure_qubits, measure_cbits, n_measures, classical_bits);
        if (res.size() > cap) fail(WQC_TN_INVALID_ARG, "output capacity too small for the distinct outcomes");
        for (size_t i = 0; i < res.size(); ++i) {
            keys[i] = res[i].key;
            probs[i] = res[i].p;
        }
        *n_out = res.size();
    });
}

namespace {

// Slices as one batch: every slice's gate list is scheduled into layers of site-disjoint
// operations up front (without running them), then step t gathers layer t of every slice into
// one one-qubit launch and shared two-site batches (merge, Jacobi rounds, factors launched once
// for units of many slices of different widths and bond caps). One executor state owns the
// stream, staging and workspace; the slices are site containers. Units of a step are cut into
// batches of at most kSliceBatchUnits units and about kSliceBatchBytes of split workspace.
// Returns false (nothing reported) if a batch fails, so that the caller reruns the slices one by
// one for per-slice statuses.
constexpr size_t kSliceBatchUnits = 512;
constexpr size_t kSliceBatchBytes = size_t(16) << 30;

bool run_slices_batched(wqc_tn_device* dev, const wqc_tn_slice* slices, size_t n_slices, const wqc_tn_gate* gates,
                        int precision, double* out_amps, int* out_status, std::string& first_error) {
    auto ex = std::make_unique<wqc_tn_state>();
    ex->init(dev, 1, 1, precision);
    DevBuf one;
    one.ensure(sizeof(double2));
    HIPC(hipMemcpy(one.p, &wqc_tn_state::kOne, sizeof(double2), hipMemcpyHostToDevice));
    std::vector<std::unique_ptr<wqc_tn_state>> st(n_slices);
    std::vector<wqc::CopyJob> starts;
    size_t max_layers = 0;
    for (size_t i = 0; i < n_slices; ++i) {
        const wqc_tn_slice& sl = slices[i];
        out_amps[2 * i] = out_amps[2 * i + 1] = std::nan("");
        out_status[i] = guarded([&] {
            auto s = std::make_unique<wqc_tn_state>();
            s->init_container(dev, sl.qubit_count, sl.max_bond_dim, precision, starts,
                              reinterpret_cast<const double*>(one.p));
            s->defer = true;
            s->queue_circuit(gates + sl.gate_offset, sl.gate_count);
            s->flush();
            max_layers = std::max(max_layers, s->layers.size());
            st[i] = std::move(s);
        });
        if (out_status[i] != WQC_TN_OK && first_error.empty())
            first_error = "slice " + std::to_string(i) + ": " + g_last_error;
    }
    // Every site starts as (1, 0): one batched copy of the complex one.
    for (size_t k = 0; k < starts.size(); k += 65536) {
        ex->tables_reset();
        std::vector<wqc::CopyJob> part(starts.begin() + k, starts.begin() + std::min(starts.size(), k + 65536));
        wqc::launch_copy_batch(ex->upload(part), (int)part.size(), 2, ex->stream);
        ex->check_launch();
        ex->sync();
    }
    try {
        for (size_t t = 0; t < max_layers; ++t) {
            std::vector<PendingOne> ones;
            std::vector<Unit> units;
            for (size_t i = 0; i < n_slices; ++i) {
                if (!st[i] || t >= st[i]->layers.size()) continue;
                wqc_tn_state::Layer& L = st[i]->layers[t];
                for (auto op : L.ones) {
                    op.owner = st[i].get();
                    ones.push_back(op);
                }
                for (auto u : L.units) {
                    u.owner = st[i].get();
                    units.push_back(u);
                }
                L = wqc_tn_state::Layer{};
            }
            ex->run_one_qubit(ones);
            for (size_t k = 0; k < units.size();) {
                size_t e = k, bytes = 0;
                while (e < units.size() && e - k < kSliceBatchUnits) {
                    const wqc_tn_state* o = units[e].owner;
                    const size_t rows = 2 * (size_t)o->sites[units[e].p].l, cols = 2 * (size_t)o->sites[units[e].p + 1].r;
                    const size_t est = 8 * rows * cols * sizeof(double2);
                    if (e > k && bytes + est > kSliceBatchBytes) break;
                    bytes += est;
                    ++e;
                }
                std::vector<Unit> chunk(units.begin() + k, units.begin() + e);
                ex->run_units(chunk);
                k = e;
            }
        }
        std::vector<wqc_tn_state*> states(n_slices);
        for (size_t i = 0; i < n_slices; ++i) states[i] = st[i].get();
        ex->amplitudes_batch(states, out_amps);
    } catch (const Failure& f) {
        if (f.code == WQC_TN_CANCELLED) throw;
        return false;
    }
    return true;
}

}  // namespace

int wqc_tn_run_slices(wqc_tn_device* dev, const wqc_tn_slice* slices, size_t n_slices,
                      const wqc_tn_gate* gates, int precision, int workers, double* out_amps,
                      int* out_status) {
    return guarded([&] {
        if (!dev || (n_slices && (!slices || !out_amps || !out_status))) fail(WQC_TN_INVALID_ARG, "null argument");
        if (!valid_precision(precision)) fail(WQC_TN_INVALID_ARG, "unknown precision mode");
        for (size_t i = 0; i < n_slices; ++i) {
            if (slices[i].qubit_count == 0) fail(WQC_TN_INVALID_ARG, "invalid qubit count: 0");
            if (slices[i].max_bond_dim == 0) fail(WQC_TN_EXEC_FAILED, "MPS max bond dimension must be > 0");
            if (slices[i].gate_count && !gates) fail(WQC_TN_INVALID_ARG, "null argument");
        }
        if (!n_slices) return;
        HIPC(hipSetDevice(dev->ordinal));
        if (workers == 0) {
            std::string err;
            if (run_slices_batched(dev, slices, n_slices, gates, precision, out_amps, out_status, err)) {
                if (!err.empty()) fail(WQC_TN_EXEC_FAILED, err);
                return;
            }
        }
        const int nw = std::max(1, std::min(workers > 0 ? workers : 8, (int)std::min<size_t>(n_slices, 64)));
        // Largest estimated cost first (two-qubit gates x chi^3 at the slice's reachable bond).
        std::vector<size_t> order(n_slices);
        std::vector<double> cost(n_slices);
        for (size_t i = 0; i < n_slices; ++i) {
            order[i] = i;
            const double b = (double)std::min<size_t>(slices[i].max_bond_dim,
                                                      (size_t)1 << std::min<size_t>(slices[i].qubit_count / 2, 30));
            cost[i] = (double)slices[i].gate_count * b * b * b;
        }
        std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return cost[a] > cost[b]; });
        std::atomic<size_t> next{0};
        std::mutex emu;
        std::string first_error;
        auto worker = [&]() {
            (void)hipSetDevice(dev->ordinal);
            std::unique_ptr<wqc_tn_state> st;
            for (;;) {
                const size_t t = next.fetch_add(1);
                if (t >= n_slices) break;
                const size_t i = order[t];
                const wqc_tn_slice& sl = slices[i];
                out_amps[2 * i] = out_amps[2 * i + 1] = std::nan("");
                const int rc = guarded([&] {
                    if (!st) {
                        st = std::make_unique<wqc_tn_state>();
                        st->init(dev, sl.qubit_count, sl.max_bond_dim, precision);
                        // Sweep graphs are captured on one thread's stream while the other
                        // workers allocate, free and copy: capture stays off here.
                        st->use_graph = false;
                    } else {
                        st->reset(sl.qubit_count, sl.max_bond_dim);
                    }
                    st->queue_circuit(gates + sl.gate_offset, sl.gate_count);
                    auto a = st->amplitude_zero();
                    out_amps[2 * i] = a.first;
                    out_amps[2 * i + 1] = a.second;
                });
                out_status[i] = rc;
                if (rc != WQC_TN_OK) {
                    std::lock_guard<std::mutex> lk(emu);
                    if (first_error.empty()) first_error = "slice " + std::to_string(i) + ": " + g_last_error;
                    if (st) {
                        st.reset();
                    }
                }
            }
        };
        std::vector<std::thread> pool;
        for (int w = 1; w < nw; ++w) pool.emplace_back(worker);
        worker();
        for (auto& th : pool) th.join();
        if (!first_error.empty()) fail(WQC_TN_EXEC_FAILED, first_error);
    });
}

namespace {

// Automaton MPO of sum_j c_j P_j for Pauli strings in Qiskit order (site q = character n-1-q).
// The bond at cut c (left of site q = c) is 2 + the number of terms whose support spans the
// cut: state 0 = nothing placed yet, 1 = term complete, 2 + k = k-th spanning term in progress.
// W[q] is D[q] x D[q+1] x 2 x 2 row-major; the coefficient sits on the term's first site.
void pauli_sum_mpo(int n, const char* labels, const double* coeffs, size_t nt,
                   std::vector<std::vector<double2>>& W, std::vector<int>& D) {
    std::vector<int> f(nt, 0), l(nt, 0);
    for (size_t j = 0; j < nt; ++j) {
        int lo = -1, hi = -1;
        for (int q = 0; q < n; ++q)
            if (labels[j * (size_t)n + (n - 1 - q)] != 'I') {
                if (lo < 0) lo = q;
                hi = q;
            }
        f[j] = lo < 0 ? 0 : lo;
        l[j] = lo < 0 ? 0 : hi;
    }
    std::vector<std::map<size_t, int>> pos(n + 1);
    D.assign(n + 1, 2);
    for (int c = 0; c <= n; ++c)
        for (size_t j = 0; j < nt; ++j)
            if (f[j] < c && c <= l[j]) pos[c][j] = D[c]++;
    auto pauli = [](char ch, double2 m[4]) {
        for (int i = 0; i < 4; ++i) m[i] = make_double2(0, 0);
        switch (ch) {
            case 'X': m[1] = m[2] = make_double2(1, 0); break;
            case 'Y': m[1] = make_double2(0, -1); m[2] = make_double2(0, 1); break;
            case 'Z': m[0] = make_double2(1, 0); m[3] = make_double2(-1, 0); break;
            default: m[0] = m[3] = make_double2(1, 0); break;
        }
    };
    W.assign(n, {});
    for (int q = 0; q < n; ++q) {
        const int Dl = D[q], Dr = D[q + 1];
        auto& w = W[q];
        w.assign((size_t)Dl * Dr * 4, make_double2(0, 0));
        auto add = [&](int a, int b, const double2 m[4], double c) {
            for (int i = 0; i < 4; ++i) {
                w[((size_t)a * Dr + b) * 4 + i].x += c * m[i].x;
                w[((size_t)a * Dr + b) * 4 + i].y += c * m[i].y;
            }
        };
        double2 id[4];
        pauli('I', id);
        add(0, 0, id, 1.0);
        add(1, 1, id, 1.0);
        for (size_t j = 0; j < nt; ++j) {
            double2 m[4];
            pauli(labels[j * (size_t)n + (n - 1 - q)], m);
            if (f[j] == q && l[j] == q) add(0, 1, m, coeffs[j]);
            else if (f[j] == q && l[j] > q) add(0, pos[q + 1][j], m, coeffs[j]);
            else if (f[j] < q && q < l[j]) add(pos[q][j], pos[q + 1][j], m, 1.0);
            else if (f[j] < q && q == l[j]) add(pos[q][j], 1, m, 1.0);
        }
    }
}

void mat2_mul(const double2 a[4], const double2 b[4], double2 c[4]) {
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) {
            double2 acc = make_double2(0, 0);
            for (int k = 0; k < 2; ++k) {
                const double2 x = a[2 * i + k], y = b[2 * k + j];
                acc.x += x.x * y.x - x.y * y.y;
                acc.y += x.x * y.y + x.y * y.x;
            }
            c[2 * i + j] = acc;
        }
}

void mat2_adj(const double2 a[4], double2 c[4]) {
    c[0] = make_double2(a[0].x, -a[0].y);
    c[1] = make_double2(a[2].x, -a[2].y);
    c[2] = make_double2(a[1].x, -a[1].y);
    c[3] = make_double2(a[3].x, -a[3].y);
}

bool is_rotation(int k) { return k == WQC_TN_GATE_RX || k == WQC_TN_GATE_RY || k == WQC_TN_GATE_RZ; }
bool is_noop(int k) { return k == WQC_TN_GATE_MEASURE || k == WQC_TN_GATE_RESET || k == WQC_TN_GATE_IF; }

}  // namespace

int wqc_tn_backward(wqc_tn_device* dev, size_t qubit_count, size_t max_bond_dim, int precision,
                    const wqc_tn_gate* gates, size_t count, const char* labels, const double* coeffs,
                    size_t n_terms, size_t checkpoint_every, size_t adjoint_bond, double* grad,
                    double* value) {
    return guarded([&] {
        if (!dev || !grad || !value || (count && !gates) || !labels || !coeffs) fail(WQC_TN_INVALID_ARG, "null argument");
        if (qubit_count == 0) fail(WQC_TN_INVALID_ARG, "invalid qubit count: 0");
        if (max_bond_dim == 0) fail(WQC_TN_EXEC_FAILED, "MPS max bond dimension must be > 0");
        if (!n_terms) fail(WQC_TN_INVALID_ARG, "observable must have at least one term");
        if (!valid_precision(precision)) fail(WQC_TN_INVALID_ARG, "unknown precision mode");
        (void)adjoint_bond;  // no adjoint state is carried: the sweep is over site cotangents
        const int n = (int)qubit_count;
        for (size_t j = 0; j < n_terms * qubit_count; ++j) {
            const char ch = labels[j];
            if (ch != 'I' && ch != 'X' && ch != 'Y' && ch != 'Z')
                fail(WQC_TN_INVALID_ARG, "invalid pauli character in label " + std::to_string(j / qubit_count));
        }
        HIPC(hipSetDevice(dev->ordinal));
        // Runs of one-qubit gates and the multi-qubit blocks after them: run r = [rb[r], re[r]),
        // block r = [re[r], rb[r+1]) (no-op gates are skipped where they stand).
        std::vector<size_t> rb, re;
        {
            size_t i = 0;
            while (i <= count) {
                rb.push_back(i);
                while (i < count && (is_one_qubit(gates[i].kind) || is_noop(gates[i].kind))) ++i;
                re.push_back(i);
                while (i < count && !is_one_qubit(gates[i].kind)) ++i;
                if (i >= count) break;
            }
        }
        const size_t R = rb.size();
        const size_t k = checkpoint_every ? checkpoint_every : 1;
        // Segment s: runs [s k, (s + 1) k) and their blocks; its checkpoint is the state at its start.
        const size_t nseg = (R + k - 1) / k;
        auto seg_begin = [&](size_t sg) { return rb[sg * k]; };
        auto seg_end = [&](size_t sg) { return (sg + 1) * k < R ? rb[(sg + 1) * k] : count; };
        // The differentiated forward: plain FP64 exact splits, one gate per flush, so that a
        // segment's replay repeats it bit for bit and its tape is in execution order.
        auto phi = std::make_unique<wqc_tn_state>();
        phi->init(dev, qubit_count, max_bond_dim, WQC_TN_PRECISION_FP64);
        phi->exact_only = true;
        phi->use_lq = false;
        auto apply = [&](size_t b, size_t e) {
            for (size_t i = b; i < e; ++i) {
                if (is_noop(gates[i].kind)) continue;
                if (phi->tape) phi->tape->tag = (int)i;
                phi->queue_gate(gates[i]);
                phi->flush();
            }
        };
        std::vector<wqc_tn_state::Snapshot> ck(nseg);
        struct Guard {
            std::vector<wqc_tn_state::Snapshot>& v;
            ~Guard() {
                for (auto& s : v) s.release();
            }
        } guard{ck};
        for (size_t sg = 0; sg < nseg; ++sg) {
            phi->snapshot(ck[sg]);
            apply(seg_begin(sg), seg_end(sg));
        }
        // Value: sum_j c_j <psi|P_j|psi>.
        {
            std::vector<double> vr(n_terms), vi(n_terms);
            phi->pauli_expectations(labels, n_terms, vr.data(), vi.data());
            double a = 0, b = 0;
            for (size_t j = 0; j < n_terms; ++j) {
                a += coeffs[j] * vr[j];
                b += coeffs[j] * vi[j];
            }
            value[0] = a;
            value[1] = b;
        }
        // Cotangents of the final sites, 2 d Re <lambda|psi> / d site with lambda = H psi (the
        // Pauli sum as an MPO times psi).
        std::vector<std::unique_ptr<DevBuf>> cot;
        {
            std::vector<std::vector<double2>> W;
            std::vector<int> D;
            pauli_sum_mpo(n, labels, coeffs, n_terms, W, D);
            size_t dmax = 1;
            for (int d : D) dmax = std::max(dmax, (size_t)d);
            auto lam = std::make_unique<wqc_tn_state>();
            lam->init(dev, qubit_count, dmax * max_bond_dim, WQC_TN_PRECISION_FP64);
            lam->set_mpo_product(*phi, W, D);
            phi->site_cotangents(*lam, cot, 2.0);
        }
        for (size_t i = 0; i < count; ++i) grad[i] = 0.0;
        // Reverse mode over the site tensors, segment by segment from the last: each segment is
        // replayed from its checkpoint with the tape on, then walked back. A one-qubit update U
        // on site q passes the cotangent back as U^H cot_q and undoes itself on the site; a
        // rotation exp(-i t P / 2) adds dE/dt = Im sum_{s,t} P[s,t] M[s,t] / 2 with
        // M[s,t] = sum conj(cot_q[a,s,b]) site_q[a,t,b]. A unit passes (cot_p, cot_{p+1}) back
        // through its truncated split (the adjoint of W = U S V^H with the factors U_k sqrt(S),
        // sqrt(S) V_k^H, including the terms between kept and discarded singular vectors and
        // the projections onto the null spaces), its gate and its merge, and puts its input
        // sites back.
        DevBuf bws;    // one unit's adjoint workspace
        DevBuf corr;   // the 2x2 correlations of a segment's rotations
        auto cot_ptr = [&](int q) { return (double2*)cot[q]->p; };
        for (size_t sgi = nseg; sgi-- > 0;) {
            phi->restore(ck[sgi]);
            Tape tape;
            phi->tape = &tape;
            apply(seg_begin(sgi), seg_end(sgi));
            phi->tape = nullptr;
            phi->sync();
            size_t nrot = 0;
            for (const auto& t : tape.ones)
                if (t.tag >= 0 && is_rotation(gates[t.tag].kind)) ++nrot;
            corr.ensure(std::max<size_t>(nrot, 1) * 4 * sizeof(double2));
            std::vector<int> rot_tag;
            for (size_t e = tape.order.size(); e-- > 0;) {
                phi->sync();
                phi->tables_reset();
                const auto ent = tape.order[e];
                if (ent.first == 0) {
                    const TapeOne& t = tape.ones[ent.second];
                    Site& st = phi->sites[t.site];
                    if (t.tag >= 0 && is_rotation(gates[t.tag].kind)) {
                        std::vector<wqc::SiteCorrJob> cj{wqc::SiteCorrJob{
                            cot_ptr(t.site), st.ptr, (double2*)corr.p + 4 * rot_tag.size(), st.l, st.r}};
                        wqc::launch_site_corr(phi->upload(cj), 1, phi->stream);
                        phi->check_launch();
                        rot_tag.push_back(t.tag);
                    }
                    double2 ua[4];
                    mat2_adj(t.u, ua);
                    std::vector<OneQubitJob> oj(2);
                    oj[0].site = cot_ptr(t.site);
                    oj[1].site = st.ptr;
                    for (auto& j : oj) {
                        j.l = st.l;
                        j.r = st.r;
                        for (int x = 0; x < 4; ++x) j.u[x] = ua[x];
                    }
                    wqc::launch_one_qubit(phi->upload(oj), 2, st.l * st.r, phi->stream);
                    phi->check_launch();
                    continue;
                }
                const TapeUnit& t = tape.units[ent.second];
                if (!t.done) fail(WQC_TN_EXEC_FAILED, "adjoint replay: a unit's split was not recorded");
                const int pp = t.pp, qq = t.qq, c = t.c, dl = t.dl, dr = t.dr, bond = t.bond;
                const size_t nJu = (size_t)pp * c, nC = (size_t)pp * pp, nY = (size_t)pp * qq;
                const size_t nL = (size_t)2 * dl * bond, nR = (size_t)bond * 2 * dr;
                bws.ensure((3 * nJu + nC + 3 * nY + nL + nR + 64) * sizeof(double2));
                double2* Ju = (double2*)bws.p;
                double2* XV = Ju + nJu;
                double2* Ec = XV + nJu;
                double2* Cm = Ec + nJu;
                double2* Y = Cm + nC;
                double2* MW = Y + nY;
                double2* TH = MW + nY;
                double2* Lb = TH + nY;
                double2* Rbar = Lb + nL;
                double2* cP = cot_ptr(t.p);      // (dl, 2, c)
                double2* cQ = cot_ptr(t.p + 1);  // (c, 2, dr)
                using WS = wqc_tn_state;
                // JU = J Abar_W, XV = X Bbar_W^H; for a tall theta' (W = theta'^H) the factors
                // swap roles: Abar_W = cot_{p+1}^H, Bbar_W = cot_p^H.
                std::vector<GemmJob> g1;
                if (!t.tall) {
                    g1.push_back(WS::gj(t.J, pp, 0, cP, c, 0, Ju, c, pp, c, pp, 1.0, 0.0));
                    g1.push_back(WS::gj(t.X, qq, 0, cQ, qq, 1, XV, c, pp, c, qq, 1.0, 0.0));
                } else {
                    g1.push_back(WS::gj(t.J, pp, 0, cQ, pp, 1, Ju, c, pp, c, pp, 1.0, 0.0));
                    g1.push_back(WS::gj(t.X, qq, 0, cP, c, 0, XV, c, pp, c, qq, 1.0, 0.0));
                }
                phi->gemm(g1, 5);
                std::vector<int> inv(pp, -1);
                for (int i = 0; i < c; ++i) inv[t.krow[i]] = i;
                wqc::SvdAdjJob aj;
                aj.JU = Ju;
                aj.XV = XV;
                aj.sig = phi->upload(t.sig);
                aj.inv = phi->upload(inv);
                aj.C = Cm;
                aj.E = Ec;
                aj.p = pp;
                aj.c = c;
                aj.floor2 = NULL_ROWS_SQ / pp;
                wqc::launch_svd_adj_coef(aj, phi->stream);
                phi->check_launch();
                // Wbar = J^H (C X + E Bbar_W).
                std::vector<GemmJob> g2{WS::gj(Cm, pp, 0, t.X, qq, 0, Y, qq, pp, qq, pp, 1.0, 0.0)};
                phi->gemm(g2, 5);
                std::vector<GemmJob> g3{!t.tall ? WS::gj(Ec, c, 0, cQ, qq, 0, Y, qq, pp, qq, c, 1.0, 1.0)
                                                : WS::gj(Ec, c, 0, cP, c, 1, Y, qq, pp, qq, c, 1.0, 1.0)};
                phi->gemm(g3, 5);
                std::vector<GemmJob> g4{WS::gj(t.J, pp, 1, Y, qq, 0, MW, qq, pp, qq, pp, 1.0, 0.0)};
                phi->gemm(g4, 5);
                // theta'bar (2dl x 2dr), then back through the gate and the merge theta = L R.
                double2* thb = MW;
                if (t.tall) {
                    wqc::launch_conj_transpose(MW, TH, pp, qq, phi->stream);
                    thb = TH;
                }
                wqc::launch_apply4_adj(thb, dl, dr, t.u4, phi->stream);
                phi->check_launch();
                std::vector<GemmJob> g5{WS::gj(thb, 2 * dr, 0, t.R, 2 * dr, 1, Lb, bond, 2 * dl, bond, 2 * dr, 1.0, 0.0),
                                        WS::gj(t.L, bond, 1, thb, 2 * dr, 0, Rbar, 2 * dr, bond, 2 * dr, 2 * dl, 1.0, 0.0)};
                phi->gemm(g5, 5);
                // Growing a cotangent or site buffer frees the old one: nothing queued may still
                // read it.
                phi->sync();
                cot[t.p]->ensure(nL * sizeof(double2));
                cot[t.p + 1]->ensure(nR * sizeof(double2));
                HIPC(hipMemcpyAsync(cot_ptr(t.p), Lb, nL * sizeof(double2), hipMemcpyDeviceToDevice, phi->stream));
                HIPC(hipMemcpyAsync(cot_ptr(t.p + 1), Rbar, nR * sizeof(double2), hipMemcpyDeviceToDevice, phi->stream));
                phi->ensure_site_cap(t.p, nL);
                phi->ensure_site_cap(t.p + 1, nR);
                HIPC(hipMemcpyAsync(phi->sites[t.p].ptr, t.L, nL * sizeof(double2), hipMemcpyDeviceToDevice, phi->stream));
                HIPC(hipMemcpyAsync(phi->sites[t.p + 1].ptr, t.R, nR * sizeof(double2), hipMemcpyDeviceToDevice,
                                    phi->stream));
                phi->sites[t.p].l = dl;
                phi->sites[t.p].r = bond;
                phi->sites[t.p + 1].l = bond;
                phi->sites[t.p + 1].r = dr;
            }
            phi->sync();
            if (!rot_tag.empty()) {
                std::vector<double2> m(4 * rot_tag.size());
                HIPC(hipMemcpy(m.data(), corr.p, m.size() * sizeof(double2), hipMemcpyDeviceToHost));
                for (size_t x = 0; x < rot_tag.size(); ++x) {
                    const int kind = gates[rot_tag[x]].kind;
                    const double2* M = &m[4 * x];
                    double v = 0.0;  // Im sum P[s,t] M[s,t]
                    if (kind == WQC_TN_GATE_RX) v = M[1].y + M[2].y;
                    if (kind == WQC_TN_GATE_RY) v = -M[1].x + M[2].x;  // P = [[0, -i], [i, 0]]
                    if (kind == WQC_TN_GATE_RZ) v = M[0].y - M[3].y;
                    grad[rot_tag[x]] += 0.5 * v;
                }
            }
        }
    });
}

int wqc_tn_contract_to_statevector(wqc_tn_state* st, double* out, size_t len_complex) {
    return guarded([&] {
        if (!st || !out) fail(WQC_TN_INVALID_ARG, "null argument");
        st->statevector(out, len_complex);
    });
}

int wqc_tn_bond_dims(wqc_tn_state* st, size_t* out, size_t len) {
    return guarded([&] {
        if (!st || !out) fail(WQC_TN_INVALID_ARG, "null argument");
        st->flush();
        for (int i = 0; i < st->n && (size_t)i < len; ++i) out[i] = (size_t)st->sites[i].r;
    });
}

int wqc_tn_download_site(wqc_tn_state* st, size_t site, double* out, size_t len_complex, size_t* l,
                         size_t* r) {
    return guarded([&] {
        if (!st || site >= (size_t)st->n) fail(WQC_TN_INVALID_ARG, "bad site");
        st->flush();
        st->sync();
        const Site& s = st->sites[site];
        if (l) *l = (size_t)s.l;
        if (r) *r = (size_t)s.r;
        const size_t cnt = (size_t)s.l * 2 * s.r;
        if (out && len_complex >= cnt)
            HIPC(hipMemcpy(out, s.ptr, cnt * sizeof(double2), hipMemcpyDeviceToHost));
    });
}

int wqc_tn_upload_site(wqc_tn_state* st, size_t site, const double* data, size_t l, size_t r) {
    return guarded([&] {
        if (!st || !data || site >= (size_t)st->n || l == 0 || r == 0) fail(WQC_TN_INVALID_ARG, "bad site upload");
        if (l > (1u << 16) || r > (1u << 16)) fail(WQC_TN_INVALID_ARG, "bond too large");
        st->flush();
        st->sync();
        st->ensure_site_cap((int)site, l * 2 * r);
        HIPC(hipMemcpy(st->sites[site].ptr, data, l * 2 * r * sizeof(double2), hipMemcpyHostToDevice));
        st->sites[site].l = (int)l;
        st->sites[site].r = (int)r;
    });
}

int wqc_tn_copy_sites(wqc_tn_state* dst, const wqc_tn_state* src) {
    return guarded([&] {
        if (!dst || !src) fail(WQC_TN_INVALID_ARG, "null state");
        if (dst == src) return;
        if (dst->dev != src->dev) fail(WQC_TN_INVALID_ARG, "states on different devices");
        if (dst->n != src->n) fail(WQC_TN_INVALID_ARG, "mismatched register");
        HIPC(hipSetDevice(dst->dev->ordinal));
        dst->flush();
        HIPC(hipStreamSynchronize(src->stream));
        dst->sync();
        for (int i = 0; i < dst->n; ++i) {
            const auto& s = src->sites[i];
            const size_t elems = (size_t)s.l * 2 * s.r;
            dst->ensure_site_cap(i, elems);
            HIPC(hipMemcpyAsync(dst->sites[i].ptr, s.ptr, elems * sizeof(double2), hipMemcpyDeviceToDevice, dst->stream));
            dst->sites[i].l = s.l;
            dst->sites[i].r = s.r;
        }
        dst->sync();
    });
}

const char* wqc_tn_backend_label(const wqc_tn_state*) { return "hip"; }

uint64_t wqc_tn_peak_vram_bytes(const wqc_tn_state* st) { return st ? st->stats.peak_vram_bytes : 0; }

int wqc_tn_get_stats(const wqc_tn_state* st, wqc_tn_stats* out) {
    if (!st || !out) {
        g_last_error = "null argument";
        return WQC_TN_INVALID_ARG;
    }
    return guarded([&] {
        const_cast<wqc_tn_state*>(st)->read_jstats();
        *out = st->stats;
    });
}

int wqc_tn_kernel_times(const wqc_tn_state* st, double* ms, size_t len) {
    if (!st || !ms) return 0;
    size_t k = std::min<size_t>(len, 8);
    for (size_t i = 0; i < k; ++i) ms[i] = st->kms[i];
    return (int)k;
}

int wqc_tn_enable_timing(wqc_tn_state* st, int enable) {
    if (!st) return WQC_TN_INVALID_ARG;
    st->timing = enable != 0;
    return WQC_TN_OK;
}

int wqc_tn_kernel_launches(const wqc_tn_state* st, uint64_t* count, size_t len) {
    if (!st || !count) return 0;
    size_t k = std::min<size_t>(len, 8);
    for (size_t i = 0; i < k; ++i) count[i] = st->klaunch[i];
    return (int)k;
}

int wqc_tn_enable_split_log(wqc_tn_state* st, int enable) {
    if (!st) return WQC_TN_INVALID_ARG;
    st->split_log_on = enable != 0;
    if (!enable) {
        st->split_log.clear();
        st->split_dw.clear();
        st->split_probe.clear();
    }
    return WQC_TN_OK;
}

size_t wqc_tn_split_log(const wqc_tn_state* st, int32_t* out, size_t cap_entries) {
    if (!st) return 0;
    const size_t m = st->split_log.size() / 8;
    if (out) {
        const size_t k = std::min(m, cap_entries);
        std::memcpy(out, st->split_log.data(), k * 8 * sizeof(int32_t));
    }
    return m;
}

size_t wqc_tn_split_discarded(const wqc_tn_state* st, double* out, size_t cap_entries) {
    if (!st) return 0;
    const size_t m = st->split_dw.size();
    if (out) std::memcpy(out, st->split_dw.data(), std::min(m, cap_entries) * sizeof(double));
    return m;
}

size_t wqc_tn_split_probe_residuals(const wqc_tn_state* st, double* out, size_t cap_entries) {
    if (!st) return 0;
    const size_t m = st->split_probe.size();
    if (out) std::memcpy(out, st->split_probe.data(), std::min(m, cap_entries) * sizeof(double));
    return m;
}

}  // extern "C"

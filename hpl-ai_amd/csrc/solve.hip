// solve.hip — host orchestration of the whole HPL-AI solve on one GPU.
//
// Replaces the driver's timed region (main.cpp:379-433): RDMAPanelLU::run
// (panel_trf.hpp:703-856) followed by iterative_refinement
// (iterative_refinement.hpp:18-57).  The whole matrix lives in HBM as FP32
// (column-major, lda = n rounded to 64), so the 2-D block-cyclic grid and its
// broadcasts become in-device hand-offs through the quantized panel buffers.
//
// LU schedule, two levels:
//  - inner block b (the reference's b): diagonal LU, FP32 TRSMs, quantization;
//  - outer block B = outer * b: the trailing matrix is updated once per outer
//    step with K = B, which divides its HBM traffic by `outer`.  With outer = 1
//    this is exactly the reference's right-looking schedule.
// Per outer step (columns c0 .. c0+B):
//   1. panel: for each inner block t: block column t (inside the outer block
//      and below it) and then the rest of block row t inside the outer block
//      take the earlier inner blocks at once (left-looking, K = t b); getrf,
//      L := A U^-1 below the diagonal block, quantize L; U := L^-1 A for the
//      rest of block row t inside the outer block, quantize;
//   2. block row: for each t: row block t takes the earlier row blocks at
//      once (left-looking, K = t b), then U12_t := L_tt^-1 A12_t, quantize;
//   3. trailing: A22 -= Lq(:, 0..B) Uq(0..B, :) (K = B).
// The left-looking parts write each block column / row block of the outer
// block once instead of once per inner block, which cuts their HBM traffic.
// With lookahead, step 3 first updates the next outer panel's columns, then
// the panel factorization of step k+1 runs on a second stream while the rest
// of step 3 runs on the main stream; with the row look-ahead (lookahead 2)
// step 3 next updates the rows of block row k+1, and step 2 of k+1 follows
// the panel on the second stream while the main stream updates the rows below.
#include <chrono>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "internal.hpp"

using namespace hplai;

namespace {

struct Buffers {
    float *A = nullptr;
    double *d = nullptr, *rhs = nullptr, *x = nullptr, *r = nullptr, *work = nullptr, *norms = nullptr;
    void *Lq[2] = {nullptr, nullptr}, *Uq[2] = {nullptr, nullptr};
    int32_t *eL[2] = {nullptr, nullptr}, *eU[2] = {nullptr, nullptr};
    float *amaxL[2] = {nullptr, nullptr}, *amaxU[2] = {nullptr, nullptr};
    int *info = nullptr;
    double *chk = nullptr;
    float *inv = nullptr;   // diagonal sub-block inverses, one set per diagonal block
    double *tsc = nullptr;  // one block of the IR's triangular solves (scratch)
    uint16_t *P = nullptr;  // FP16 trailing storage (n x n, leading dimension n)
    uint16_t *Q = nullptr;  // its back buffer (n x n), when on
};

void free_buffers(Buffers &B, bool keep_factors)
{
    auto f = [](void *p) {
        if (p) (void)hipFree(p);
    };
    if (!keep_factors) {
        f(B.A);
        f(B.x);
    }
    f(B.d);
    f(B.rhs);
    f(B.r);
    f(B.tsc);
    f(B.work);
    f(B.norms);
    f(B.info);
    f(B.chk);
    f(B.inv);
    f(B.P);
    f(B.Q);
    for (int i = 0; i < 2; ++i) {
        f(B.Lq[i]);
        f(B.Uq[i]);
        f(B.eL[i]);
        f(B.eU[i]);
        f(B.amaxL[i]);
        f(B.amaxU[i]);
    }
}

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Per-launch HIP events for the Schur updates (roofline accounting) and per
// phase of the reference's Timer (timer.hpp:58-237); only when profiling.
struct GemmTimer {
    bool on = false;
    std::vector<hipEvent_t> ev;   // start/stop pairs
    std::vector<double> flops, bytes;
    std::vector<char> large, inner, cls;
    std::vector<hipEvent_t> pev;  // phase start/stop pairs
    std::vector<int> pitem;
    std::vector<double> pops;
    int rc = HPLAI_OK;
    void pbegin(hipStream_t s)
    {
        if (!on || rc) return;
        hipEvent_t a = nullptr, b = nullptr;
        if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
            rc = HPLAI_EHIP;
            return;
        }
        pev.push_back(a);
        pev.push_back(b);
        (void)hipEventRecord(a, s);
    }
    void pend(hipStream_t s, int item, double ops)
    {
        if (!on || rc || pev.empty()) return;
        (void)hipEventRecord(pev.back(), s);
        pitem.push_back(item);
        pops.push_back(ops);
    }
    void begin(hipStream_t s, double f, double by, bool is_large, bool is_inner, int c)
    {
        if (!on || rc) return;
        hipEvent_t a = nullptr, b = nullptr;
        if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
            rc = HPLAI_EHIP;
            return;
        }
        ev.push_back(a);
        ev.push_back(b);
        flops.push_back(f);
        bytes.push_back(by);
        large.push_back(is_large ? 1 : 0);
        inner.push_back(is_inner ? 1 : 0);
        cls.push_back(static_cast<char>(c));
        (void)hipEventRecord(a, s);
    }
    void end(hipStream_t s)
    {
        if (!on || rc || ev.empty()) return;
        (void)hipEventRecord(ev.back(), s);
    }
    // after synchronisation
    // after synchronisation; recomputes every field, so it can be called again
    void collect(hplai_result *r)
    {
        for (int i = 0; i < HPLAI_PHASES; ++i) r->phase_seconds[i] = r->phase_ops[i] = 0.0;
        double ms = 0.0, fl = 0.0, lms = 0.0, lfl = 0.0, lby = 0.0, ims = 0.0, ifl = 0.0;
        int64_t ln = 0, in = 0;
        for (int c = 0; c < 4; ++c) {
            r->schur_class_launches[c] = 0;
            r->schur_class_flops[c] = r->schur_class_ms[c] = 0.0;
        }
        for (size_t i = 0; i + 1 < ev.size(); i += 2) {
            float t = 0.0f;
            if (hipEventElapsedTime(&t, ev[i], ev[i + 1]) != hipSuccess) t = 0.0f;
            ms += t;
            fl += flops[i / 2];
            const int c = cls[i / 2];
            r->schur_class_launches[c] += 1;
            r->schur_class_flops[c] += flops[i / 2];
            r->schur_class_ms[c] += t;
            if (large[i / 2]) {
                ++ln;
                lms += t;
                lfl += flops[i / 2];
                lby += bytes[i / 2];
            }
            if (inner[i / 2]) {
                ++in;
                ims += t;
                ifl += flops[i / 2];
            }
        }
        r->schur_launches = static_cast<int64_t>(flops.size());
        r->schur_flops = fl;
        r->schur_ms = ms;
        r->schur_large_launches = ln;
        r->schur_large_flops = lfl;
        r->schur_large_ms = lms;
        r->schur_large_bytes = lby;
        r->schur_inner_launches = in;
        r->schur_inner_flops = ifl;
        r->schur_inner_ms = ims;
        for (size_t i = 0; i < ev.size() / 2; ++i) {
            float t = 0.0f;
            if (hipEventElapsedTime(&t, ev[2 * i], ev[2 * i + 1]) != hipSuccess) t = 0.0f;
            r->phase_seconds[HPLAI_PHASE_GEMM_UPDATE] += 1e-3 * t;
            r->phase_ops[HPLAI_PHASE_GEMM_UPDATE] += flops[i];
        }
        for (size_t i = 0; i < pitem.size(); ++i) {
            float t = 0.0f;
            if (hipEventElapsedTime(&t, pev[2 * i], pev[2 * i + 1]) != hipSuccess) t = 0.0f;
            r->phase_seconds[pitem[i]] += 1e-3 * t;
            r->phase_ops[pitem[i]] += pops[i];
        }
    }
    ~GemmTimer()
    {
        for (auto e : ev) (void)hipEventDestroy(e);
        for (auto e : pev) (void)hipEventDestroy(e);
    }
};

// "!epoch" progress lines (panel_trf.hpp:730-745): an event on the main stream
// at the start of every outer step; a line is printed once its event has
// completed (checked without blocking at each step, all at the end), with
// the elapsed GPU time since the LU started.  One GPU is a 1 x 1 grid, so the
// reference's epoch is one panel of b columns; lines come at outer steps.
struct EpochLog {
    bool on = false;
    int64_t n = 0;
    int b = 1;
    hipEvent_t t0 = nullptr;
    std::vector<hipEvent_t> ev;
    std::vector<int64_t> col;
    size_t printed = 0;
    void start(hipStream_t s)
    {
        if (on && hipEventCreate(&t0) == hipSuccess) (void)hipEventRecord(t0, s);
    }
    void mark(hipStream_t s, int64_t c)
    {
        if (!on || t0 == nullptr) return;
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return;
        (void)hipEventRecord(e, s);
        ev.push_back(e);
        col.push_back(c);
    }
    void flush(bool wait)
    {
        for (; printed < ev.size(); ++printed) {
            if (!wait && hipEventQuery(ev[printed]) != hipSuccess) break;
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, t0, ev[printed]) != hipSuccess) ms = 0.0f;
            const double el = 1e-3 * ms;
            const double dn = static_cast<double>(n), nrem = static_cast<double>(n - col[printed]);
            const double flop = 2. / 3. * (dn * dn * dn - nrem * nrem * nrem);
            std::printf("!epoch %lld/%lld: elapsed=%f, %f Pflops (estimate)\n", (long long)(col[printed] / b),
                        (long long)(n / b), el, el > 0 ? 1.e-15 * flop / el : 0.0);
            std::fflush(stdout);
        }
    }
    ~EpochLog()
    {
        if (t0) (void)hipEventDestroy(t0);
        for (auto e : ev) (void)hipEventDestroy(e);
    }
};

// classes of Schur launches (hplai_result.schur_class_*)
constexpr int kTrailing = 0, kBulk = 1, kInBlock = 2, kBlockRow = 3;

struct LU {
    int64_t n, lda, ldq;
    int b, outer, fmt, kpad, es;
    // 0, or h (0 < h < outer): the outer panel is factored in two parts; after
    // its first h inner blocks the rest of the outer block (and of its block
    // row) takes them in one update (K = h b), and the later inner blocks
    // look left only as far as block h
    int split = 0;
    bool lazy;
    bool higham = false;          // Higham's matrix instead of the LCG matrix
    bool packed = false;          // Lq / Uq in the panel-packed layout (b % 256 == 0)
    bool trail16 = false;         // trailing updates into the FP16 storage B->P (lazy only)
    float pscale = 1.0f;          // B->P holds pscale times the partial sums
    float pdown = 1.0f;           // (float)(1 / pscale), the lazy initialisation's scale
    bool bback = false;           // back buffer B->Q of the FP16 trailing storage
    double h_alpha = 0.0, h_beta = 0.0;
    uint64_t seed;
    int64_t ntiles;
    Buffers *B;
    hipStream_t main, side, bulk;
    hipEvent_t *ev;
    int num_cu;
    GemmTimer *timer;
    EpochLog *epochs = nullptr;

    bool tiled() const { return fmt != HPLAI_FMT_F16; }
    float *a(int64_t i, int64_t j) const { return B->A + j * lda + i; }
    // slice t of row (column) `row` of the quantized panels; packed: row is a
    // multiple of 256 and the K offset moves whole 16 KB stage blocks
    int64_t qoff(int64_t row, int t) const
    {
        const int64_t kb = static_cast<int64_t>(t) * kpad * es;
        return row * ldq * es + (packed ? 256 * kb : kb);
    }
    uint8_t *lq(int buf, int64_t row, int t) const { return static_cast<uint8_t *>(B->Lq[buf]) + qoff(row, t); }
    uint8_t *uq(int buf, int64_t col, int t) const { return static_cast<uint8_t *>(B->Uq[buf]) + qoff(col, t); }
    int32_t *el(int buf, int t, int64_t row) const
    {
        return tiled() ? B->eL[buf] + t * ntiles + row / b : nullptr;
    }
    int32_t *eu(int buf, int t, int64_t col) const
    {
        return tiled() ? B->eU[buf] + t * ntiles + col / b : nullptr;
    }
    float *aml(int buf, int t, int64_t row) const
    {
        return tiled() ? B->amaxL[buf] + t * ntiles + row / b : nullptr;
    }
    float *amu(int buf, int t, int64_t col) const
    {
        return tiled() ? B->amaxU[buf] + t * ntiles + col / b : nullptr;
    }
    const float *dinv(int64_t c) const { return B->inv + (c / b) * diag_inv_floats(b); }

    // cls: kTrailing, an outer step's update of blocks that have not joined a
    // panel yet (to the FP16 trailing storage when it is on); kBulk, the rows
    // below the outer block taking the earlier panels of the block; kInBlock,
    // an update inside the outer panel; kBlockRow, a row block of the block
    // row.  Each class is timed apart.
    int gemm(int buf, int t0, int nsl, int64_t row0, int64_t col0, int64_t m, int64_t nn,
             hipStream_t s, int cls) const
    {
        const bool trailing = cls == kTrailing, in_block = cls == kInBlock;
        if (m <= 0 || nn <= 0) return HPLAI_OK;
        if (trailing && trail16) {
            const double K = static_cast<double>(nsl) * b;
            timer->begin(s, 2.0 * static_cast<double>(m) * static_cast<double>(nn) * K,
                         static_cast<double>(m + nn) * K * es + 4.0 * static_cast<double>(m) * nn,
                         schur_large_tiles(fmt, m, nn, b, num_cu), false, cls);
            const int rc = launch_schur_c16(fmt, m, nn, nsl, kpad, lq(buf, row0, t0), ldq, el(buf, t0, row0),
                                            ntiles, uq(buf, col0, t0), ldq, eu(buf, t0, col0), ntiles, b,
                                            B->P + col0 * n + row0, n, pscale, num_cu, s, packed);
            timer->end(s);
            return rc;
        }
        const double K = static_cast<double>(nsl) * b;
        timer->begin(s, 2.0 * static_cast<double>(m) * static_cast<double>(nn) * K,
                     static_cast<double>(m + nn) * K * es + 8.0 * static_cast<double>(m) * nn,
                     schur_large_tiles(fmt, m, nn, b, num_cu), in_block, cls);
        const int rc = launch_schur(fmt, m, nn, nsl, kpad, lq(buf, row0, t0), ldq, el(buf, t0, row0),
                                    ntiles, uq(buf, col0, t0), ldq, eu(buf, t0, col0), ntiles, b,
                                    a(row0, col0), lda, num_cu, s, packed);
        timer->end(s);
        return rc;
    }

    // one phase of the reference's Timer around the launches of f
    template <typename F>
    int timed(int item, double ops, hipStream_t s, F &&f) const
    {
        timer->pbegin(s);
        const int rc = f();
        timer->pend(s, item, ops);
        return rc;
    }

    // Lazy initialisation (-d): add the generated a_ij to the partial sums of
    // an m x k block when it joins a panel (LazyInitializer, lazy_init.hpp:92-227).
    int lazy_init(int64_t i0, int64_t j0, int64_t m, int64_t k, hipStream_t s) const
    {
        if (!lazy || m <= 0 || k <= 0) return HPLAI_OK;
        return timed(HPLAI_PHASE_LAZY_INIT, static_cast<double>(m) * k, s, [&] {
            if (higham) return launch_gen_higham(h_alpha, h_beta, i0, j0, m, k, a(i0, j0), lda, 1, s);
            if (trail16)   // A += a_ij + P / pscale (the partial sums of earlier outer steps)
                return launch_gen_matrix(n, seed, i0, j0, m, k, B->d, a(i0, j0), lda, 1, s, B->P + j0 * n + i0,
                                         n, pdown, bback ? B->Q + j0 * n + i0 : nullptr);
            return launch_gen_matrix(n, seed, i0, j0, m, k, B->d, a(i0, j0), lda, 1, s);
        });
    }

    // Step 1 for the outer block at c0 with sw inner blocks.  Two streams:
    // s carries the chain from one diagonal block to the next (getrf, the
    // inverses, the left-looking updates, TRSMs and quantization of the rows
    // and columns inside the outer block); the rows below the outer block (their
    // left-looking update, lazy initialisation, TRSM and quantization) go to
    // the bulk stream, which waits for the chain only where it needs its
    // results.  s waits for the bulk stream at the end.
    int panel(int64_t c0, int sw, int buf, hipStream_t s, bool join = true) const
    {
        int rc = HPLAI_OK;
        const int64_t Bw = static_cast<int64_t>(sw) * b;
        const int64_t rb0 = c0 + Bw, mb = n - rb0;   // rows below the outer block
        hipStream_t q = bulk;
        hipEvent_t ev_fork = ev[3], ev_inv = ev[4], ev_u = ev[5], ev_join = ev[6];
        if (tiled()) {
            rc = launch_fill_f32(B->amaxL[buf], static_cast<int64_t>(outer) * ntiles, 0.0f, s);
            if (rc == HPLAI_OK) rc = launch_fill_f32(B->amaxU[buf], static_cast<int64_t>(outer) * ntiles, 0.0f, s);
            if (rc) return rc;
        }
        if (mb > 0) {
            HPLAI_HIP(hipEventRecord(ev_fork, s));
            HPLAI_HIP(hipStreamWaitEvent(q, ev_fork, 0));
        }
        const bool two = split > 0 && split < sw;   // the panel is factored in two parts
        for (int t = 0; t < sw && rc == HPLAI_OK; ++t) {
            const int64_t c = c0 + static_cast<int64_t>(t) * b;
            const double bd = static_cast<double>(b);
            // inner blocks base .. t - 1 are applied per block column; with the
            // split, at t == split the rest of the outer block takes blocks
            // 0 .. split - 1 in one update (mid) and later blocks start there
            const bool mid = two && t == split;
            const int base = two && t >= split ? split : 0, nsl = t - base;
            // bulk: the rows below take all earlier panels of the block at once
            // (left-looking, K = t b) once the chain has quantized U of block
            // t - 1 (the wait refers to the record of the previous iteration)
            if (mb > 0) {
                if (t > 0) {
                    HPLAI_HIP(hipStreamWaitEvent(q, ev_u, 0));
                    if (mid) rc = gemm(buf, 0, split, rb0, c, mb, rb0 - c, q, kBulk);
                    if (rc == HPLAI_OK && nsl > 0) rc = gemm(buf, base, nsl, rb0, c, mb, b, q, kBulk);
                }
                if (rc == HPLAI_OK) rc = lazy_init(rb0, c, mb, b, q);
            }
            // chain: the diagonal block and the rows of the block column inside
            // the outer block take the t earlier inner blocks at once (K = t b);
            // mid: the whole lower-right part of the outer block
            if (rc == HPLAI_OK && mid) rc = gemm(buf, 0, split, c, c, rb0 - c, rb0 - c, s, kInBlock);
            if (rc == HPLAI_OK && nsl > 0) rc = gemm(buf, base, nsl, c, c, rb0 - c, b, s, kInBlock);
            if (rc == HPLAI_OK) rc = lazy_init(c, c, rb0 - c, b, s);
            if (rc == HPLAI_OK)
                rc = timed(HPLAI_PHASE_DIAG_LU, 2.0 / 3.0 * bd * bd * bd, s, [&] {
                    int r = launch_getrf(b, a(c, c), lda, B->info + c / b, s);
                    // the inverses also serve the IR's block solves, so the
                    // last diagonal block gets them too
                    if (r == HPLAI_OK && b <= 512)
                        r = launch_diag_inv(b, a(c, c), lda, const_cast<float *>(dinv(c)), s);
                    return r;
                });
            if (rc) break;
            // bulk: L := A U^-1 and quantization of the rows below
            if (mb > 0) {
                HPLAI_HIP(hipEventRecord(ev_inv, s));
                HPLAI_HIP(hipStreamWaitEvent(q, ev_inv, 0));
                rc = timed(HPLAI_PHASE_TRSM_L, static_cast<double>(mb) * bd * bd, q, [&] {
                    return launch_trsm_l(mb, b, a(c, c), lda, a(rb0, c), lda, aml(buf, t, rb0), b,
                                         b <= 512 ? dinv(c) : nullptr, q);
                });
                if (rc == HPLAI_OK)
                    rc = timed(HPLAI_PHASE_CONV_L, static_cast<double>(mb) * bd, q, [&] {
                        return launch_quantize(fmt, 0, mb, b, b, a(rb0, c), lda, lq(buf, rb0, t), ldq,
                                               aml(buf, t, rb0), el(buf, t, rb0), q, packed);
                    });
                if (rc) break;
            }
            // chain: the same for the rows inside the outer block
            const int64_t mrow = rb0 - c - b;
            if (mrow <= 0) continue;
            rc = timed(HPLAI_PHASE_TRSM_L, static_cast<double>(mrow) * bd * bd, s, [&] {
                return launch_trsm_l(mrow, b, a(c, c), lda, a(c + b, c), lda, aml(buf, t, c + b), b,
                                     b <= 512 ? dinv(c) : nullptr, s);
            });
            if (rc == HPLAI_OK)
                rc = timed(HPLAI_PHASE_CONV_L, static_cast<double>(mrow) * bd, s, [&] {
                    return launch_quantize(fmt, 0, mrow, b, b, a(c + b, c), lda, lq(buf, c + b, t), ldq,
                                           aml(buf, t, c + b), el(buf, t, c + b), s, packed);
                });
            const int64_t ncols = rb0 - (c + b);
            if (rc) break;
            // block row t inside the outer block, likewise from the left (K = t b)
            if (nsl > 0) rc = gemm(buf, base, nsl, c, c + b, b, ncols, s, kInBlock);
            if (rc == HPLAI_OK) rc = lazy_init(c, c + b, b, ncols, s);
            if (rc == HPLAI_OK)
                rc = timed(HPLAI_PHASE_TRSM_R, static_cast<double>(ncols) * bd * bd, s, [&] {
                    return launch_trsm_u(b, ncols, a(c, c), lda, a(c, c + b), lda, amu(buf, t, c + b), b,
                                         b <= 512 ? dinv(c) : nullptr, s);
                });
            if (rc == HPLAI_OK)
                rc = timed(HPLAI_PHASE_CONV_R, static_cast<double>(ncols) * bd, s, [&] {
                    return launch_quantize(fmt, 1, ncols, b, b, a(c, c + b), lda, uq(buf, c + b, t), ldq,
                                           amu(buf, t, c + b), eu(buf, t, c + b), s, packed);
                });
            if (rc) break;
            if (mb > 0) HPLAI_HIP(hipEventRecord(ev_u, s));
        }
        if (mb > 0) {
            HPLAI_HIP(hipEventRecord(ev_join, q));
            if (join) HPLAI_HIP(hipStreamWaitEvent(s, ev_join, 0));
        }
        return rc;
    }

    // Step 2: block row U12 of the outer block at c0.
    int block_row(int64_t c0, int sw, int buf, hipStream_t s) const
    {
        int rc = HPLAI_OK;
        const int64_t Bw = static_cast<int64_t>(sw) * b;
        const int64_t col0 = c0 + Bw, ncols = n - col0;
        if (ncols <= 0) return rc;
        const bool two = split > 0 && split < sw;
        for (int t = 0; t < sw && rc == HPLAI_OK; ++t) {
            const int64_t c = c0 + static_cast<int64_t>(t) * b;
            const double bd = static_cast<double>(b);
            const int base = two && t >= split ? split : 0, nsl = t - base;
            // with the split, row blocks split .. sw - 1 take blocks 0 .. split - 1 at once
            if (two && t == split) rc = gemm(buf, 0, split, c, col0, Bw - static_cast<int64_t>(split) * b, ncols, s,
                                             kBlockRow);
            // row block t takes the earlier row blocks at once (left-looking, K = t b)
            if (rc == HPLAI_OK && nsl > 0) rc = gemm(buf, base, nsl, c, col0, b, ncols, s, kBlockRow);
            if (rc == HPLAI_OK) rc = lazy_init(c, col0, b, ncols, s);   // block row right of the outer panel
            if (rc == HPLAI_OK)
                rc = timed(HPLAI_PHASE_TRSM_R, static_cast<double>(ncols) * bd * bd, s, [&] {
                    return launch_trsm_u(b, ncols, a(c, c), lda, a(c, col0), lda, amu(buf, t, col0), b,
                                         b <= 512 ? dinv(c) : nullptr, s);
                });
            if (rc == HPLAI_OK)
                rc = timed(HPLAI_PHASE_CONV_R, static_cast<double>(ncols) * bd, s, [&] {
                    return launch_quantize(fmt, 1, ncols, b, b, a(c, col0), lda, uq(buf, col0, t), ldq,
                                           amu(buf, t, col0), eu(buf, t, col0), s, packed);
                });
        }
        return rc;
    }

    // lookahead 0: update, then the next panel, all on the main stream.
    // lookahead 1: the next outer panel's columns first, then the panel on the
    // side stream while the main stream updates the rest.
    // lookahead 2: as 1, and the rest is updated in two parts: the rows of
    // the next block row first, so that the side stream computes that block
    // row (after the panel chain) while the main stream updates the rows below.
    int run(int lookahead) const
    {
        int rc = HPLAI_OK;
        int64_t c0 = 0;
        int step = 0;
        int sw = static_cast<int>(std::min<int64_t>(outer, (n - c0) / b));
        bool row_done = false;   // block row of the current step computed ahead
        if (epochs) epochs->start(main);
        rc = panel(c0, sw, 0, main);
        while (rc == HPLAI_OK && c0 < n) {
            const int buf = step & 1;
            if (epochs) {
                epochs->mark(main, c0);
                epochs->flush(false);
            }
            const int64_t Bw = static_cast<int64_t>(sw) * b;
            if (!row_done) rc = block_row(c0, sw, buf, main);
            row_done = false;
            if (rc) break;
            const int64_t r0 = c0 + Bw, rest = n - r0;
            const int64_t c1 = r0;
            const int sw1 = rest > 0 ? static_cast<int>(std::min<int64_t>(outer, rest / b)) : 0;
            if (rest > 0) {
                if (lookahead) {
                    // next panel's columns first, then hand them to the side stream
                    const int64_t B1 = static_cast<int64_t>(sw1) * b;
                    const bool ahead = lookahead > 1 && rest > B1;
                    rc = gemm(buf, 0, sw, r0, r0, rest, B1, main, kTrailing);
                    if (rc) break;
                    HPLAI_HIP(hipEventRecord(ev[0], main));
                    HPLAI_HIP(hipStreamWaitEvent(side, ev[0], 0));
                    // with the row look-ahead the main stream joins the panel's
                    // bulk stream itself (the block row does not need its rows)
                    rc = panel(c1, sw1, buf ^ 1, side, !ahead);
                    if (rc) break;
                    if (ahead) {
                        rc = gemm(buf, 0, sw, r0, r0 + B1, B1, rest - B1, main, kTrailing);
                        if (rc) break;
                        HPLAI_HIP(hipEventRecord(ev[2], main));
                        HPLAI_HIP(hipStreamWaitEvent(side, ev[2], 0));
                        rc = block_row(c1, sw1, buf ^ 1, side);
                        if (rc) break;
                        HPLAI_HIP(hipEventRecord(ev[1], side));
                        rc = gemm(buf, 0, sw, r0 + B1, r0 + B1, rest - B1, rest - B1, main, kTrailing);
                        if (rc) break;
                        HPLAI_HIP(hipStreamWaitEvent(main, ev[6], 0));   // the panel's bulk stream
                        row_done = true;
                    } else {
                        HPLAI_HIP(hipEventRecord(ev[1], side));
                        rc = gemm(buf, 0, sw, r0, r0 + B1, rest, rest - B1, main, kTrailing);
                        if (rc) break;
                    }
                    HPLAI_HIP(hipStreamWaitEvent(main, ev[1], 0));
                } else {
                    rc = gemm(buf, 0, sw, r0, r0, rest, rest, main, kTrailing);
                    if (rc == HPLAI_OK) rc = panel(c1, sw1, buf ^ 1, main);
                    if (rc) break;
                }
            }
            // back-buffer write-back every 32 trailing updates (BackBuffer::write_back,
            // back_buffer.hpp:81-170, BB_NCYCLE = 32), over the blocks that do not
            // join the next outer block (those merge P and Q at their lazy initialisation)
            if (bback && (step + 1) % 32 == 0 && rest > 0) {
                const int64_t r1 = c1 + static_cast<int64_t>(sw1) * b;
                rc = launch_backbuffer(n - r1, n - r1, B->P + r1 * n + r1, B->Q + r1 * n + r1, n, main);
                if (rc) break;
            }
            c0 = c1;
            sw = sw1;
            ++step;
        }
        return rc;
    }
};

// Fills the LU descriptor for an n x n matrix in B->A (leading dimension lda).
void init_lu(LU &lu, hplai_ctx ctx, int64_t n, int64_t lda, int b, int fmt, int outer, bool lazy,
             uint64_t seed, Buffers *B, GemmTimer *timer)
{
    const int es = fmt == HPLAI_FMT_E4M3 ? 1 : 2;
    const int bk = 128 / es;
    lu.timer = timer;
    lu.n = n;
    lu.lda = lda;
    lu.b = b;
    lu.outer = outer;
    lu.fmt = fmt;
    lu.es = es;
    lu.kpad = (b + bk - 1) / bk * bk;
    lu.ldq = static_cast<int64_t>(outer) * lu.kpad;
    // every panel row / column offset is a multiple of b, so with b % 256 == 0
    // the operands can use the panel-packed layout (whole 1 KB wave loads)
    lu.packed = b % 256 == 0;
    lu.ntiles = n / b;
    lu.lazy = lazy;
    lu.seed = seed;
    lu.B = B;
    lu.main = nullptr;
    lu.side = ctx->side;
    lu.bulk = ctx->bulk;
    lu.ev = ctx->ev;
    lu.num_cu = ctx->num_cu;
}

// Allocates the LU's working buffers: quantized panel double buffers, tile
// exponents and maxima, diagonal inverses, pivot flags.  On failure the
// buffers allocated so far stay in B for free_buffers.
int alloc_lu_buffers(const LU &lu, Buffers &B, hipStream_t s)
{
    auto alloc = [](void **p, size_t bytes, const char *what) {
        const hipError_t e = hipMalloc(p, bytes);
        return e == hipSuccess ? HPLAI_OK : hip_fail(e, what);
    };
    const int64_t nt = lu.ntiles;
    int rc = alloc(reinterpret_cast<void **>(&B.info), sizeof(int) * nt, "hipMalloc info");
    if (rc == HPLAI_OK)
        rc = alloc(reinterpret_cast<void **>(&B.inv), sizeof(float) * diag_inv_floats(lu.b) * nt, "hipMalloc inv");
    for (int i = 0; i < 2 && rc == HPLAI_OK; ++i) {
        const size_t qbytes = static_cast<size_t>(lu.n) * lu.ldq * lu.es;
        rc = alloc(&B.Lq[i], qbytes, "hipMalloc Lq");
        if (rc == HPLAI_OK) rc = alloc(&B.Uq[i], qbytes, "hipMalloc Uq");
        if (rc == HPLAI_OK && hipMemsetAsync(B.Lq[i], 0, qbytes, s) != hipSuccess) rc = HPLAI_EHIP;
        if (rc == HPLAI_OK && hipMemsetAsync(B.Uq[i], 0, qbytes, s) != hipSuccess) rc = HPLAI_EHIP;
        if (rc == HPLAI_OK && lu.fmt != HPLAI_FMT_F16) {
            const size_t cnt = static_cast<size_t>(lu.outer) * nt;
            rc = alloc(reinterpret_cast<void **>(&B.eL[i]), sizeof(int32_t) * cnt, "hipMalloc eL");
            if (rc == HPLAI_OK) rc = alloc(reinterpret_cast<void **>(&B.eU[i]), sizeof(int32_t) * cnt, "hipMalloc eU");
            if (rc == HPLAI_OK) rc = alloc(reinterpret_cast<void **>(&B.amaxL[i]), sizeof(float) * cnt, "hipMalloc amaxL");
            if (rc == HPLAI_OK) rc = alloc(reinterpret_cast<void **>(&B.amaxU[i]), sizeof(float) * cnt, "hipMalloc amaxU");
        }
    }
    if (rc == HPLAI_OK && hipMemsetAsync(B.info, 0, sizeof(int) * nt, s) != hipSuccess) rc = HPLAI_EHIP;
    return rc;
}

// After synchronisation: HPLAI_EPIVOT with the first zero / non-finite pivot.
int check_pivots(const LU &lu, const Buffers &B, const char *who)
{
    std::vector<int> info(lu.ntiles);
    HPLAI_HIP(hipMemcpy(info.data(), B.info, sizeof(int) * lu.ntiles, hipMemcpyDeviceToHost));
    for (int64_t q = 0; q < lu.ntiles; ++q)
        if (info[q] != 0) {
            set_error("%s: zero or non-finite pivot at column %lld", who,
                      (long long)(q * lu.b + info[q] - 1));
            return HPLAI_EPIVOT;
        }
    return HPLAI_OK;
}

int check_lu_args(int64_t n, int b, int fmt, const char *who)
{
    HPLAI_CHECK(n > 0 && b > 0 && n % b == 0, "%s: need n %% b == 0 (n=%lld b=%d)", who, (long long)n, b);
    HPLAI_CHECK(b <= 1024, "%s: b=%d > 1024", who, b);
    HPLAI_CHECK(fmt == HPLAI_FMT_F16 || fmt == HPLAI_FMT_F16_TILE || fmt == HPLAI_FMT_E4M3,
                "%s: unknown fmt %d", who, fmt);
    if (fmt != HPLAI_FMT_F16)
        HPLAI_CHECK(b % 128 == 0, "%s: tile-scaled formats need b %% 128 == 0 (b=%d)", who, b);
    return HPLAI_OK;
}

}  // namespace

extern "C" {

int hplai_factor_f32(hplai_ctx ctx, int64_t n, int b, int fmt, int outer, int lookahead, float *A,
                     int64_t lda)
{
    HPLAI_CHECK(ctx != nullptr && A != nullptr, "hplai_factor_f32: NULL argument");
    HPLAI_CHECK(lda >= n, "hplai_factor_f32: lda < n");
    int rc = check_lu_args(n, b, fmt, "hplai_factor_f32");
    if (rc) return rc;
    HPLAI_HIP(hipSetDevice(ctx->device));
    LU lu;
    Buffers B;
    GemmTimer timer;
    init_lu(lu, ctx, n, lda, b, fmt, outer > 0 ? outer : 1, false, 0, &B, &timer);
    B.A = A;   // the caller's matrix: not freed below
    rc = alloc_lu_buffers(lu, B, nullptr);
    if (rc == HPLAI_OK) rc = lu.run(lookahead);
    if (hipDeviceSynchronize() != hipSuccess && rc == HPLAI_OK) rc = hip_fail(hipGetLastError(), "hplai_factor_f32");
    if (rc == HPLAI_OK) rc = check_pivots(lu, B, "hplai_factor_f32");
    B.A = nullptr;
    free_buffers(B, false);
    return rc;
}

int hplai_lu_solve_f64(hplai_ctx ctx, int trans, int64_t n, int nrhs, int bs, const float *LU,
                       int64_t lda, double *X, int64_t ldx, void *stream)
{
    HPLAI_CHECK(ctx != nullptr && LU != nullptr && X != nullptr, "hplai_lu_solve_f64: NULL argument");
    HPLAI_CHECK(n >= 0 && nrhs >= 0 && bs > 0 && bs <= 1024 && lda >= n && (nrhs <= 1 || ldx >= n),
                "hplai_lu_solve_f64: bad sizes (n=%lld nrhs=%d bs=%d lda=%lld ldx=%lld)", (long long)n,
                nrhs, bs, (long long)lda, (long long)ldx);
    HPLAI_CHECK(nrhs <= 65535, "hplai_lu_solve_f64: nrhs %d > 65535", nrhs);
    if (n == 0 || nrhs == 0) return HPLAI_OK;
    HPLAI_HIP(hipSetDevice(ctx->device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    int rc;
    if (trans == 0) {
        rc = launch_trsv_lower(n, bs, LU, lda, X, s, nrhs, ldx);
        if (rc == HPLAI_OK) rc = launch_trsv_upper(n, bs, LU, lda, X, s, nrhs, ldx);
    } else {
        rc = launch_trsv_upper_t(n, bs, LU, lda, X, s, nrhs, ldx);
        if (rc == HPLAI_OK) rc = launch_trsv_lower_t(n, bs, LU, lda, X, s, nrhs, ldx);
    }
    return rc;
}

int hplai_release(hplai_ctx ctx)
{
    if (ctx == nullptr) return HPLAI_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->A) (void)hipFree(ctx->A);
    if (ctx->x) (void)hipFree(ctx->x);
    ctx->A = nullptr;
    ctx->x = nullptr;
    ctx->n = ctx->lda = 0;
    return HPLAI_OK;
}

int hplai_solution_x(hplai_ctx ctx, double *x_host)
{
    HPLAI_CHECK(ctx != nullptr && ctx->x != nullptr && x_host != nullptr,
                "solution_x: no kept solution (run hplai_solve with keep_factors)");
    HPLAI_HIP(hipSetDevice(ctx->device));
    HPLAI_HIP(hipMemcpy(x_host, ctx->x, sizeof(double) * ctx->n, hipMemcpyDeviceToHost));
    return HPLAI_OK;
}

int hplai_solution_factors(hplai_ctx ctx, float *lu_host, int64_t ldh)
{
    HPLAI_CHECK(ctx != nullptr && ctx->A != nullptr && lu_host != nullptr && ldh >= ctx->n,
                "solution_factors: no kept factors or ldh < n");
    HPLAI_HIP(hipSetDevice(ctx->device));
    HPLAI_HIP(hipMemcpy2D(lu_host, sizeof(float) * ldh, ctx->A, sizeof(float) * ctx->lda,
                          sizeof(float) * ctx->n, ctx->n, hipMemcpyDeviceToHost));
    return HPLAI_OK;
}

int hplai_solution_block(hplai_ctx ctx, int64_t i0, int64_t j0, int64_t m, int64_t k, float *host,
                         int64_t ldh)
{
    HPLAI_CHECK(ctx != nullptr && ctx->A != nullptr && host != nullptr,
                "solution_block: no kept factors (run hplai_solve with keep_factors)");
    HPLAI_CHECK(i0 >= 0 && j0 >= 0 && m >= 0 && k >= 0 && i0 + m <= ctx->n && j0 + k <= ctx->n && ldh >= m,
                "solution_block: block (%lld, %lld) of %lld x %lld outside the %lld x %lld factors or ldh < m",
                (long long)i0, (long long)j0, (long long)m, (long long)k, (long long)ctx->n, (long long)ctx->n);
    if (m == 0 || k == 0) return HPLAI_OK;
    HPLAI_HIP(hipSetDevice(ctx->device));
    HPLAI_HIP(hipMemcpy2D(host, sizeof(float) * ldh, ctx->A + j0 * ctx->lda + i0, sizeof(float) * ctx->lda,
                          sizeof(float) * m, k, hipMemcpyDeviceToHost));
    return HPLAI_OK;
}

int hplai_solve(hplai_ctx ctx, const hplai_params *p, hplai_result *res)
{
    HPLAI_CHECK(ctx != nullptr && p != nullptr && res != nullptr, "hplai_solve: NULL argument");
    HPLAI_HIP(hipSetDevice(ctx->device));
    memset(res, 0, sizeof(*res));
    const int64_t n = p->n;
    const int b = p->b;
    const int fmt = p->fmt;
    const int outer = p->outer > 0 ? p->outer : 1;
    const int maxit = p->maxit > 0 ? (p->maxit < 64 ? p->maxit : 64) : 50;
    int rc = check_lu_args(n, b, fmt, "hplai_solve");
    if (rc) return rc;
    hplai_release(ctx);

    LU lu;
    Buffers B;
    GemmTimer timer;
    timer.on = p->profile != 0;
    init_lu(lu, ctx, n, (n + 63) / 64 * 64, b, fmt, outer, p->lazy != 0, p->seed, &B, &timer);
    HPLAI_CHECK(p->matrix == HPLAI_MATRIX_HPLAI || p->matrix == HPLAI_MATRIX_HIGHAM ||
                    p->matrix == HPLAI_MATRIX_HIGHAM_KAPPA,
                "hplai_solve: unknown matrix family %d", p->matrix);
    if (p->matrix != HPLAI_MATRIX_HPLAI) {
        // HMGen(n, 1000., 0.5, d), main.cpp:258-259: the reference's signs
        // (+alpha, +beta), or Fasi and Higham's (-alpha, -beta) with cond = kappa
        lu.higham = true;
        rc = higham_coeffs(n, p->cond > 0.0 ? p->cond : 1000.0, 0.5, &lu.h_alpha, &lu.h_beta);
        if (rc) return rc;
        if (p->matrix == HPLAI_MATRIX_HIGHAM_KAPPA) {
            lu.h_alpha = -lu.h_alpha;
            lu.h_beta = -lu.h_beta;
        }
    }
    HPLAI_CHECK(p->split >= 0 && (p->split == 0 || p->split < outer),
                "hplai_solve: split %d must be 0 or below outer %d", p->split, outer);
    lu.split = p->split;
    if (p->trail16) {
        // 16-bit trailing storage (the reference's -d double-decker lower deck):
        // partial sums of not yet factored blocks in FP16, scaled by
        // scalea = sqrt(n sqrt(n)) (hpl_rand.hpp:154) and merged by the lazy
        // initialisation with downscale = 1 / scalea (panel_trf.hpp:343-345)
        HPLAI_CHECK(p->trail16 == 1 || p->trail16 == 2, "hplai_solve: trail16 must be 0, 1 or 2");
        HPLAI_CHECK(p->lazy != 0, "hplai_solve: the FP16 trailing storage needs lazy initialisation");
        HPLAI_CHECK(p->matrix == HPLAI_MATRIX_HPLAI, "hplai_solve: the FP16 trailing storage takes the HPL-AI matrix");
        HPLAI_CHECK(fmt != HPLAI_FMT_F16_TILE && b % 128 == 0,
                    "hplai_solve: the FP16 trailing storage needs b %% 128 == 0 and the FP16 or e4m3 format");
        lu.trail16 = true;
        lu.bback = p->trail16 == 2;
        const double sa = std::sqrt(static_cast<double>(n) * std::sqrt(static_cast<double>(n)));
        lu.pscale = static_cast<float>(sa);
        lu.pdown = static_cast<float>(1.0 / static_cast<double>(lu.pscale));
    }
    EpochLog epochs;
    epochs.on = p->verbose != 0;
    epochs.n = n;
    epochs.b = b;
    lu.epochs = &epochs;
    hipStream_t s = nullptr;

#define ALLOC(ptr, bytes)                                                             \
    do {                                                                              \
        hipError_t _e = hipMalloc(reinterpret_cast<void **>(&(ptr)), (bytes));        \
        if (_e != hipSuccess) {                                                       \
            rc = hip_fail(_e, "hplai_solve: hipMalloc " #ptr);                        \
            free_buffers(B, false);                                                   \
            return rc;                                                                \
        }                                                                             \
    } while (0)
    ALLOC(B.A, sizeof(float) * lu.lda * n);
    ALLOC(B.d, sizeof(double) * n);
    ALLOC(B.rhs, sizeof(double) * n);
    ALLOC(B.x, sizeof(double) * n);
    ALLOC(B.r, sizeof(double) * n);
    ALLOC(B.tsc, sizeof(double) * b);
    ALLOC(B.work, sizeof(double) * std::max(otf_gemv_work_doubles(n), higham_gemv_work_doubles(n)));
    ALLOC(B.norms, sizeof(double) * 4);
    ALLOC(B.chk, sizeof(double) * 3);
    if (lu.trail16) ALLOC(B.P, sizeof(uint16_t) * n * n);
    if (lu.bback) ALLOC(B.Q, sizeof(uint16_t) * n * n);
#undef ALLOC
    rc = alloc_lu_buffers(lu, B, s);
    if (rc) {
        free_buffers(B, false);
        return rc;
    }

    // Everything after the allocations runs in one body whose every exit
    // (HPLAI_HIP returns from it) comes back here, so the buffers are freed
    // on every error path.
    bool converged = false;
    int steps = 0;
    auto body = [&]() -> int {
        // ---- generation (main.cpp:256-275), outside the timed region ----
        HPLAI_HIP(hipDeviceSynchronize());
        double t0 = now_s();
        if (rc == HPLAI_OK)
            rc = lu.higham ? launch_higham_diag(lu.h_alpha, lu.h_beta, 0, n, B.d, s)   // pdiaggen(HMGen)
                           : launch_gen_diag(n, p->seed, 0, n, B.d, s);
        if (lu.lazy) {
            // pmatgen0 (matgen.hpp:80-105): partial sums start at zero
            HPLAI_HIP(hipMemsetAsync(B.A, 0, sizeof(float) * lu.lda * n, s));
            if (lu.trail16) HPLAI_HIP(hipMemsetAsync(B.P, 0, sizeof(uint16_t) * n * n, s));
            if (lu.bback) HPLAI_HIP(hipMemsetAsync(B.Q, 0, sizeof(uint16_t) * n * n, s));
        } else if (rc == HPLAI_OK) {
            rc = lu.higham ? launch_gen_higham(lu.h_alpha, lu.h_beta, 0, 0, n, n, B.A, lu.lda, 0, s)
                           : launch_gen_matrix(n, p->seed, 0, 0, n, n, B.d, B.A, lu.lda, 0, s);
        }
        if (rc == HPLAI_OK) rc = launch_gen_rhs(n, p->seed, 0, n, B.rhs, s);
        HPLAI_HIP(hipDeviceSynchronize());
        double t1 = now_s();
        res->time_gen = t1 - t0;

        // ---- LU ----
        if (rc == HPLAI_OK) rc = lu.run(p->lookahead);
        HPLAI_HIP(hipDeviceSynchronize());
        epochs.flush(true);
        double t2 = now_s();
        res->time_lu = t2 - t1;
        if (rc == HPLAI_OK) rc = timer.rc;
        timer.collect(res);
        if (rc == HPLAI_OK) rc = check_pivots(lu, B, "hplai_solve");

        // ---- iterative refinement (iterative_refinement.hpp:18-57) ----
        if (rc == HPLAI_OK) {
            double hn[4];
            rc = launch_norm_inf(n, B.rhs, B.norms + 0, s);
            if (rc == HPLAI_OK)   // x and r are scratch until x0 is set
                rc = lu.higham ? launch_higham_norm(n, lu.h_alpha, lu.h_beta, B.x, B.r, B.work, B.norms + 1, s)
                               : launch_norm_inf(n, B.d, B.norms + 1, s);
            if (rc == HPLAI_OK) rc = launch_vec_div(n, B.rhs, B.d, B.x, s);   // x0 = b ./ diag(A)
            HPLAI_HIP(hipMemcpy(hn, B.norms, sizeof(double) * 2, hipMemcpyDeviceToHost));
            // hpl_infnorm (panel_norm.hpp:78-84): 2 max d for the LCG matrix; the
            // Higham matrix's largest row sum (higham_infnorm, panel_norm.hpp:87-135)
            const double normb = hn[0], norma = lu.higham ? hn[1] : 2.0 * hn[1];
            for (int it = 0; it < maxit && rc == HPLAI_OK; ++it) {
                HPLAI_HIP(hipMemcpyAsync(B.r, B.rhs, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
                rc = launch_norm_inf(n, B.x, B.norms + 2, s);
                const double dn2 = static_cast<double>(n) * static_cast<double>(n);
                if (rc == HPLAI_OK)
                    rc = lu.timed(HPLAI_PHASE_IR_GEMV, 2.0 * dn2, s,
                                  [&] {
                                      if (lu.higham)
                                          return launch_higham_gemv(n, lu.h_alpha, lu.h_beta, -1.0, B.x, B.r, B.work, s);
                                      return launch_otf_gemv(n, p->seed, B.d, -1.0, B.x, B.r, B.work, s);
                                  });
                if (rc == HPLAI_OK) rc = launch_norm_inf(n, B.r, B.norms + 3, s);
                if (rc) break;
                HPLAI_HIP(hipMemcpy(hn + 2, B.norms + 2, sizeof(double) * 2, hipMemcpyDeviceToHost));
                const double normx = hn[2], normr = hn[3];
                const double h = normr / (norma * normx + normb) * 1. / (static_cast<double>(n) * DBL_EPSILON / 2);
                res->trace_resid[it] = normr;
                res->trace_harness[it] = h;
                steps = it + 1;
                if (p->verbose) {
                    printf("# iterative refinement: step=%3d, residual=%20.16e hpl-harness=%f\n", it, normr, h);
                    fflush(stdout);
                }
                if (h < 16.) {
                    res->residual = normr;
                    res->harness = h;
                    converged = true;
                    break;
                }
                rc = lu.timed(HPLAI_PHASE_IR_TRSV, 2.0 * dn2, s, [&] {
                    // diagonal blocks through the inverses the LU stored, where it stored them
                    if (tri_inverse_ok(b)) {
                        const int64_t st = diag_inv_floats(b);
                        int r = launch_trsv_lower_inv(n, b, B.A, lu.lda, B.r, B.inv, st, B.tsc, s);
                        if (r == HPLAI_OK) r = launch_trsv_upper_inv(n, b, B.A, lu.lda, B.r, B.inv, st, B.tsc, s);
                        return r;
                    }
                    int r = launch_trsv_lower(n, b, B.A, lu.lda, B.r, s);
                    if (r == HPLAI_OK) r = launch_trsv_upper(n, b, B.A, lu.lda, B.r, s);
                    return r;
                });
                if (rc == HPLAI_OK) rc = launch_vec_add(n, B.r, B.x, s);
            }
        }
        HPLAI_HIP(hipDeviceSynchronize());
        double t3 = now_s();
        res->time_ir = t3 - t2;
        res->time_total = t3 - t1;
        if (p->profile) {
            timer.collect(res);   // the IR phases recorded after the LU
            res->phase_seconds[HPLAI_PHASE_TOTAL] = res->time_total;
            double rest = res->time_total;
            for (int i = 0; i < HPLAI_PHASE_TOTAL; ++i) rest -= res->phase_seconds[i];
            res->phase_seconds[HPLAI_PHASE_MISC] = rest;
        }
        res->ir_steps = steps;
        const double dn = static_cast<double>(n);
        res->gflops = (2. / 3 * dn * dn * dn + 3. / 2 * dn * dn) / res->time_total * 1e-9;
        if (!converged) {
            res->residual = -1.0;
            res->harness = -1.0;
        }
        if (rc == HPLAI_OK) {
            rc = launch_checksum(n, B.A, lu.lda, B.chk, s);
            if (rc == HPLAI_OK)
                HPLAI_HIP(hipMemcpy(res->checksum, B.chk, sizeof(double) * 3, hipMemcpyDeviceToHost));
        }
        return rc;
    };
    rc = body();

    const bool keep = p->keep_factors != 0 && rc == HPLAI_OK;
    if (keep) {
        ctx->A = B.A;
        ctx->x = B.x;
        ctx->n = n;
        ctx->lda = lu.lda;
    }
    free_buffers(B, keep);
    if (rc == HPLAI_OK && !converged) {
        set_error("hplai_solve: harness >= 16 after %d IR steps", steps);
        rc = HPLAI_ENOCONV;
    }
    return rc;
}

}  // extern "C"

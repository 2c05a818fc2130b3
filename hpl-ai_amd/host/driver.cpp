// driver.cpp — `driver.out n b P [flags]`, the reference's command line
// (main.cpp:33-131, 144-452) over the C ABI in include/hplai.h.
//
// One MI355X holds the whole matrix, so P (the process-grid rows) is accepted
// and reported but no grid is built; -r (RDMA broadcast) and -not (2-D
// block-cyclic instead of tile layout) describe communication and layout
// choices that do not exist on one device and are accepted as no-ops.
// Output lines follow the reference formats so runs diff directly:
//   "# iterative refinement: step=%3d, residual=%20.16e hpl-harness=%f"
//   "%.9f sec. %.9f GFlop/s resid = %20.15e hpl-harness = %.9f"
//   "check %22.17e %22.17e, %22.17e"   (--checksum)
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <unistd.h>
#include <ctime>

#include "hplai.h"

static void usage()
{
    std::printf(
        "HPL-AI for one MI355X\n\n"
        "usage: ./driver.out <n> <b> <P> [arguments]\n\n"
        "   n                 The matrix size.\n"
        "   b                 The block size. Requires mod(n,b) = 0.\n"
        "   P                 Process-grid rows of the reference; reported only.\n"
        "   -not, -nt         Accepted (the layout is always one column-major matrix).\n"
        "   -d, -l            Lazy initialisation: the trailing matrix starts at zero and\n"
        "                     each block is regenerated and merged when it joins a panel;\n"
        "                     the partial sums of blocks not yet factored are kept in FP16\n"
        "                     (the reference's double-decker lower deck) where b %% 128 == 0\n"
        "                     and the format is f16 or e4m3, in FP32 otherwise.\n"
        "   --trail32         With -d: keep the partial sums in FP32.\n"
        "   --back_buffer     -d with the compensated FP16 back buffer (a second FP16\n"
        "                     n x n array, folded every 32 trailing updates).\n"
        "   -r, --rdma        Accepted (no inter-process broadcast on one device).\n"
        "   -w, --warmup      One IR step only.\n"
        "   --checksum        Print the factor checksum line.\n"
        "   --highammat       Higham's HPL-AI matrix as the reference builds it instead of\n"
        "                     the random diagonally dominant matrix.\n"
        "   --highammat-kappa Higham's matrix with cond(A, inf) = 1000 exactly.\n"
        "   --fmt <f>         Schur operands: f16 (default), f16t (per-tile 2^e), e4m3.\n"
        "   --outer <s>       Panels per trailing update (default 1).\n"
        "   --lookahead <n>   0: none; 1: factor panel k+1 on a second stream (default);\n"
        "                     2: also update the next block row ahead of the trailing matrix.\n"
        "   --seed <s>        Generator seed (default 42).\n"
        "   --device <d>      HIP device (default 0).\n"
        "   --job-id <id>     Job id in the Timerdump file name (default 0).\n"
        "   --no-timerdump    Do not time the phases or write Timerdump.<jobid>.<n>.<b>.1.0.\n"
        "   --help, -h        Print this message.\n");
}

static void print_ctime(const char *msg)
{
    std::time_t t;
    std::time(&t);
    std::printf("%s%s", msg, std::ctime(&t));
}

// Timerdump file of the reference (main.cpp:470-478, Timer::show / dump_mp
// timer.hpp:167-214): one line per phase, then the Schur GEMM rate.
static void write_timerdump(bool warmup, int jobid, long long n, int b, const hplai_result &r)
{
    static const char *names[HPLAI_PHASES] = {
        "DIAG_BCAST", "LCOL_BCAST", "RROW_BCAST", "TEST", "WAIT", "DIAG_LU", "TRSM_L", "TRSM_R",
        "CONV_L", "CONV_R", "GEMM_UPDATE", "GEMM_PROGRESS", "LAZY_INIT", "WRITE_BACK", "IR_GEMV",
        "IR_GEMV_COMM", "IR_TRSV", "IR_TRSV_MV", "IR_TRSV_COMM", "TOTAL", "MISC"};
    char filename[1024];
    std::snprintf(filename, sizeof(filename),
                  warmup ? "Warmup.%06d.%09lld.%04d.%05d.%05d" : "Timerdump.%06d.%09lld.%04d.%05d.%05d", jobid,
                  n, b, 1, 0);
    FILE *fp = std::fopen(filename, "w");
    if (!fp) return;
    char host[256] = "unknown";
    (void)gethostname(host, sizeof(host) - 1);
    std::fprintf(fp, "# row=%d, col=%d, host=%s\n", 0, 0, host);
    const double total = r.phase_seconds[HPLAI_PHASE_TOTAL];
    for (int i = 0; i < HPLAI_PHASES; ++i) {
        const double t = r.phase_seconds[i];
        std::fprintf(fp, " %-12s : %e sec : %6.2f %% : %20lld : %e Gop/s\n", names[i], t,
                     total > 0 ? 100.0 * t / total : 0.0, (long long)r.phase_ops[i],
                     t > 0 ? 1e-9 * r.phase_ops[i] / t : 0.0);
    }
    const double gt = r.phase_seconds[HPLAI_PHASE_GEMM_UPDATE] + r.phase_seconds[HPLAI_PHASE_GEMM_PROGRESS];
    const double gf = r.phase_ops[HPLAI_PHASE_GEMM_UPDATE] + r.phase_ops[HPLAI_PHASE_GEMM_PROGRESS];
    std::fprintf(fp, "GEMM_TOTAL: %f Tflops\n", gt > 0 ? 1.e-12 * gf / gt : 0.0);
    std::fclose(fp);
}

int main(int argc, char **argv)
{
    for (int i = 1; i < argc; ++i)
        if (!std::strcmp(argv[i], "--help") || !std::strcmp(argv[i], "-h")) {
            usage();
            return 0;
        }
    long long n = 120;
    int b = 3, nrow = 2;
    if (argc > 1) n = std::atoll(argv[1]);
    if (argc > 2) b = std::atoi(argv[2]);
    if (argc > 3) nrow = std::atoi(argv[3]);
    hplai_params p;
    std::memset(&p, 0, sizeof(p));
    p.n = n;
    p.b = b;
    p.seed = 42;   // main.cpp:174
    p.fmt = HPLAI_FMT_F16;
    p.outer = 1;
    p.lookahead = 1;
    p.maxit = 50;
    p.verbose = 1;
    p.profile = 1;   // per-phase times for the Timerdump file
    int jobid = 0;
    bool checksum = false, warmup = false, lazy = false, rdma = false, tile = true, higham = false;
    bool higham_kappa = false, trail32 = false, back_buffer = false;
    int device = 0;
    for (int i = 4; i < argc; ++i) {
        const char *a = argv[i];
        auto next = [&](void) -> const char * { return i + 1 < argc ? argv[++i] : ""; };
        if (!std::strcmp(a, "-not") || !std::strcmp(a, "-nt") || !std::strcmp(a, "--notile")) tile = false;
        else if (!std::strcmp(a, "-d") || !std::strcmp(a, "--delayed") || !std::strcmp(a, "-l") ||
                 !std::strcmp(a, "--lazy-init")) lazy = true;
        else if (!std::strcmp(a, "-r") || !std::strcmp(a, "--rdma")) rdma = true;
        else if (!std::strcmp(a, "-w") || !std::strcmp(a, "--warmup")) warmup = true;
        else if (!std::strcmp(a, "--checksum")) checksum = true;
        else if (!std::strcmp(a, "--trail32")) trail32 = true;
        else if (!std::strcmp(a, "--back_buffer")) lazy = back_buffer = true;   // main.cpp:54
        else if (!std::strcmp(a, "--highammat")) higham = true;
        else if (!std::strcmp(a, "--highammat-kappa")) higham = higham_kappa = true;
        else if (!std::strcmp(a, "--fmt")) {
            const char *f = next();
            if (!std::strcmp(f, "f16")) p.fmt = HPLAI_FMT_F16;
            else if (!std::strcmp(f, "f16t")) p.fmt = HPLAI_FMT_F16_TILE;
            else if (!std::strcmp(f, "e4m3") || !std::strcmp(f, "fp8")) p.fmt = HPLAI_FMT_E4M3;
            else { std::fprintf(stderr, "unknown --fmt %s\n", f); return 2; }
        }
        else if (!std::strcmp(a, "--outer")) p.outer = std::atoi(next());
        else if (!std::strcmp(a, "--lookahead")) p.lookahead = std::atoi(next());
        else if (!std::strcmp(a, "--seed")) p.seed = std::strtoull(next(), nullptr, 10);
        else if (!std::strcmp(a, "--device")) device = std::atoi(next());
        else if (!std::strcmp(a, "--job-id")) jobid = std::atoi(next());
        else if (!std::strcmp(a, "--no-timerdump")) p.profile = 0;
        else std::printf("Ignoring unknown option: %s\n", a);
    }
    if (warmup) p.maxit = 1;
    p.lazy = lazy ? 1 : 0;
    p.matrix = higham_kappa ? HPLAI_MATRIX_HIGHAM_KAPPA : (higham ? HPLAI_MATRIX_HIGHAM : HPLAI_MATRIX_HPLAI);
    // -d is the reference's double-decker layout (main.cpp:52-54): FP16 partial sums
    p.trail16 = lazy && !trail32 && b % 128 == 0 && p.fmt != HPLAI_FMT_F16_TILE && !higham ? 1 : 0;
    if (back_buffer && p.trail16) p.trail16 = 2;
    p.cond = 1000.0;

    std::printf("jobid=%d\n", jobid);
    std::printf("n=%lld b=%d r=%d c=%d\n", n, b, 1, 1);
    std::printf("%s %s %s %s nopack %s noskiplu %s\n", tile ? "tile" : "2dbc",
                lazy ? "lazy" : "precomp", rdma ? "rdma" : "mpi", warmup ? "warmup" : "full",
                checksum ? "checksum" : "nocheck", higham ? "hmat" : "ddmat");
    std::printf("device=%d P=%d fmt=%s outer=%d lookahead=%d trailing=%s\n", device, nrow,
                p.fmt == HPLAI_FMT_F16 ? "f16" : (p.fmt == HPLAI_FMT_F16_TILE ? "f16t" : "e4m3"),
                p.outer, p.lookahead, p.trail16 == 2 ? "fp16+bb" : (p.trail16 ? "fp16" : "fp32"));
    std::printf("epoch_size = %d\n", b);   // b * lcm(1, 1) on a 1 x 1 grid (main.cpp:193)
    print_ctime("#BEGIN: ");
    std::fflush(stdout);

    hplai_ctx ctx = nullptr;
    int rc = hplai_create(device, &ctx);
    if (rc != HPLAI_OK) {
        std::fprintf(stderr, "hplai_create: %s\n", hplai_last_error());
        return 1;
    }
    hplai_result r;
    rc = hplai_solve(ctx, &p, &r);
    if (rc != HPLAI_OK && rc != HPLAI_ENOCONV) {
        std::fprintf(stderr, "hplai_solve failed (%d): %s\n", rc, hplai_last_error());
        hplai_destroy(ctx);
        return 1;
    }
    if (checksum) std::printf("check %22.17e %22.17e, %22.17e\n", r.checksum[0], r.checksum[1], r.checksum[2]);
    print_ctime("#END__: ");
    std::printf("%.9f sec. %.9f GFlop/s resid = %20.15e hpl-harness = %.9f\n", r.time_total, r.gflops,
                r.residual, r.harness);
    std::printf("# generation %.6f s, LU %.6f s, IR %.6f s (%d residual evaluations)\n", r.time_gen,
                r.time_lu, r.time_ir, r.ir_steps);
    std::fflush(stdout);
    if (p.profile) write_timerdump(warmup, jobid, n, b, r);
    hplai_destroy(ctx);
    return rc == HPLAI_OK ? 0 : 3;
}

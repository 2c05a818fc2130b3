/*
 * seam_example.c — the reference's Schur seam bound to libhplai_gfx950.so
 * the way a maintainer of the reference's host code would bind it, in plain
 * C built with gcc (no hipcc, no C++).
 *
 * The reference's overload (schur_updator.hpp:28-31)
 *     inline void gemmschur(int n, int m, int k, fp16 const *a, int lda,
 *                           fp16 const *b, int ldb, float *c, int ldc)
 * keeps its signature and forwards to hplai_gemmschur_f16; its operands stay
 * in the reference's column-major layout (a: n x k L panel from
 * convert_left_panels, b: k x m U panel from convert_right_panels), now in
 * device memory.
 *
 * Test driver: seam_example IN OUT reads int32 [n, m, k, lda, ldb, ldc],
 * then a (lda*k binary16), b (ldb*m binary16), c (ldc*m FP32) from IN, runs
 * the seam on device copies and writes c to OUT.
 */
#include <hip/hip_runtime_api.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "hplai.h"

typedef uint16_t fp16;   /* the reference's fp16 is a 16-bit binary16 value */

static hplai_ctx g_hplai;   /* created once at start-up, one per rank thread */

/* schur_updator.hpp:28-31, C -= A * B (alpha = -1, beta = 1 by convention) */
static void gemmschur(int n, int m, int k, fp16 const *a, int lda, fp16 const *b, int ldb, float *c, int ldc)
{
    int rc = hplai_gemmschur_f16(g_hplai, n, m, k, a, lda, b, ldb, c, ldc, /*stream*/ NULL);
    if (rc != HPLAI_OK) {
        fprintf(stderr, "gemmschur: %s\n", hplai_last_error());
        abort();
    }
}

static int fail(const char *what)
{
    fprintf(stderr, "seam_example: %s\n", what);
    return 1;
}

int main(int argc, char **argv)
{
    if (argc != 3) return fail("usage: seam_example IN OUT");
    FILE *f = fopen(argv[1], "rb");
    if (!f) return fail("cannot open input");
    int32_t h[6];
    if (fread(h, sizeof(int32_t), 6, f) != 6) return fail("short header");
    const int n = h[0], m = h[1], k = h[2], lda = h[3], ldb = h[4], ldc = h[5];
    if (n < 0 || m < 0 || k < 0 || lda < n || ldb < k || ldc < n) return fail("bad shape");
    const size_t na = (size_t)lda * k, nb = (size_t)ldb * m, nc = (size_t)ldc * m;
    fp16 *a = malloc(na * sizeof(fp16) + 1), *b = malloc(nb * sizeof(fp16) + 1);
    float *c = malloc(nc * sizeof(float) + 1);
    if (!a || !b || !c) return fail("host allocation");
    if (fread(a, sizeof(fp16), na, f) != na || fread(b, sizeof(fp16), nb, f) != nb ||
        fread(c, sizeof(float), nc, f) != nc)
        return fail("short input");
    fclose(f);

    if (hplai_create(0, &g_hplai) != HPLAI_OK) return fail(hplai_last_error());
    fp16 *da = NULL, *db = NULL;
    float *dc = NULL;
    if (hipMalloc((void **)&da, na * sizeof(fp16) + 1) != hipSuccess ||
        hipMalloc((void **)&db, nb * sizeof(fp16) + 1) != hipSuccess ||
        hipMalloc((void **)&dc, nc * sizeof(float) + 1) != hipSuccess)
        return fail("hipMalloc");
    if (hipMemcpy(da, a, na * sizeof(fp16), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(db, b, nb * sizeof(fp16), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dc, c, nc * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
        return fail("hipMemcpy to device");

    gemmschur(n, m, k, da, lda, db, ldb, dc, ldc);

    if (hipDeviceSynchronize() != hipSuccess) return fail("hipDeviceSynchronize");
    if (hipMemcpy(c, dc, nc * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) return fail("hipMemcpy to host");
    (void)hipFree(da);
    (void)hipFree(db);
    (void)hipFree(dc);
    hplai_destroy(g_hplai);

    f = fopen(argv[2], "wb");
    if (!f || fwrite(c, sizeof(float), nc, f) != nc) return fail("cannot write output");
    fclose(f);
    free(a);
    free(b);
    free(c);
    return 0;
}

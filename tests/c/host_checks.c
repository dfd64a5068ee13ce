/* Driver for the AddressSanitizer / UndefinedBehaviorSanitizer build of the
 * library's host code (no GPU needed): the integer layout functions over
 * every (bw, m, l, row) up to bw = 64 plus spot values up to bw = 1024, and
 * every drop-in and batched entry point with bad arguments, which must be
 * rejected -- with the error recorded -- before any device work. */
#include <stdio.h>
#include <string.h>

#include "s2kit/FST_semi_fly.h"
#include "s2kit/FST_semi_memo.h"
#include "s2kit/cospml.h"
#include "s2kit/naive.h"
#include "s2kit/pml.h"
#include "s2kit/seminaive.h"
#include "s2kit/util.h"
#include "s2kit/weights.h"
#include "s2kit_amd.h"

static int fails = 0;

static void expect(int cond, const char* what) {
    if (!cond) {
        printf("FAILED: %s (last error: %s)\n", what, s2amd_last_error());
        fails++;
    }
}

static void expect_err(const char* needle, const char* what) {
    expect(strstr(s2amd_last_error(), needle) != NULL, what);
}

int main(void) {
    long sum = 0;
    for (int bw = 1; bw <= 64; ++bw) {
        for (int m = 0; m < bw; ++m) {
            sum += TableSize(m, bw);
            for (int l = m; l < bw; ++l) sum += TableOffset(m, l) + RowSize(m, l);
            for (int i = 0; i <= bw; ++i) sum += Transpose_RowSize(i, m, bw);
        }
        for (int m = 0; m <= bw; ++m) sum += Reduced_SpharmonicTableSize(bw, m) + Reduced_Naive_TableSize(bw, m);
        sum += Spharmonic_TableSize(bw);
        for (int m = -(bw - 1); m < bw; ++m)
            for (int l = m < 0 ? -m : m; l < bw; ++l) sum += IndexOfHarmonicCoeff(m, l, bw);
    }
    for (int bw = 128; bw <= 1024; bw *= 2) sum += Spharmonic_TableSize(bw) + TableSize(bw / 2, bw);
    expect(sum > 0, "layout functions");
    expect(Reduced_SpharmonicTableSize(256, 256) == 2812672, "T(256)");

    s2amd_set_abort_on_error(0);
    static _Alignas(64) double z[64];
    double* pz[8] = {z, z, z, z, z, z, z, z};
    FSTSemiMemo(z, z, z, z, 0, pz, NULL, COMPLEX, 0, NULL, NULL, z);
    expect_err("bw must be in [2, 1024]", "FSTSemiMemo bw");
    FSTSemiMemo(z, z, z, z, 8, pz, NULL, (DataFormat)3, 8, NULL, NULL, z);
    expect_err("data_format", "FSTSemiMemo data_format");
    FSTSemiMemo(z, z, z, z, 8, pz, NULL, COMPLEX, 9, NULL, NULL, z);
    expect_err("cutoff", "FSTSemiMemo cutoff");
    InvFSTSemiMemo(z, z, z, z, 8, NULL, NULL, COMPLEX, 8, NULL, NULL);
    expect_err("transpose_seminaive_naive_table is NULL", "InvFSTSemiMemo table");
    FZTSemiMemo(z, z, NULL, z, 8, z, NULL, COMPLEX, NULL, z);
    expect_err("NULL", "FZTSemiMemo rres");
    ConvOn2SphereSemiMemo(z, z, z, NULL, z, z, 8, NULL);
    expect_err("NULL", "Conv filter");
    TransMult(z, z, z, z, z, z, 2000);
    expect_err("bw", "TransMult bw");
    GenerateCosPmlTable(8, 8, z, NULL);
    expect_err("m must be in [0, bw)", "GenerateCosPmlTable m");
    GeneratePmlTable(8, -1, z, NULL);
    expect_err("m must be in [0, bw)", "GeneratePmlTable m");
    DLTSemi(z, 8, 9, z, NULL, z, z, NULL);
    expect_err("m must be in [0, bw)", "DLTSemi m");
    InvDLTNaive(NULL, 8, 0, z, z);
    expect_err("NULL", "InvDLTNaive coeffs");
    FSTSemiFly(z, z, z, z, 8, NULL, COMPLEX, 9, NULL, NULL, z);
    expect_err("cutoff", "FSTSemiFly cutoff");
    expect(SemiNaive_Naive_Pml_Table(8, 9, z, NULL) == NULL, "table builder cutoff");
    expect_err("cutoff", "table builder cutoff message");
    expect(Spharmonic_Pml_Table(8, NULL, NULL) == NULL, "table builder NULL");

    s2amd_plan* p = NULL;
    expect(s2amd_plan_create(&p, 1, 1, 4, S2AMD_FP64) == S2AMD_EINVAL && p == NULL, "plan bw");
    expect(s2amd_plan_create(&p, 16, 17, 4, S2AMD_FP64) == S2AMD_EINVAL, "plan cutoff");
    expect(s2amd_plan_create(&p, 16, 16, 0, S2AMD_FP64) == S2AMD_EINVAL, "plan fields");
    expect(s2amd_plan_create(&p, 16, 16, 4, 9) == S2AMD_EINVAL, "plan precision");
    expect(s2amd_plan_create(NULL, 16, 16, 4, 0) == S2AMD_EINVAL, "plan out");
    expect(s2amd_forward(NULL, z, z, z, z, 1, 0, NULL) == S2AMD_EINVAL, "forward plan");
    expect(s2amd_forward_noalloc(NULL, z, z, z, z, 1, 0, NULL) == S2AMD_EINVAL, "forward_noalloc plan");
    expect(s2amd_inverse(NULL, z, z, z, z, 1, 0, NULL) == S2AMD_EINVAL, "inverse plan");
    expect(s2amd_plan_destroy(NULL) == 0, "destroy NULL");
    s2amd_conv_request rq = {8, z, z, z, z, z, NULL};
    s2amd_conv_batch* b = NULL;
    expect(s2amd_conv_batch_create(&b, &rq, 1, S2AMD_FP64, NULL) == S2AMD_EINVAL, "conv batch NULL grid");
    expect(s2amd_conv_batch_create(&b, &rq, 0, S2AMD_FP64, NULL) == S2AMD_EINVAL, "conv batch n");
    expect(s2amd_conv_batch_run(NULL, NULL) == S2AMD_EINVAL, "conv batch run");
    expect(s2amd_conv_batch_destroy(NULL) == 0, "conv batch destroy NULL");
    expect(strcmp(s2amd_stage_name(0), "") != 0, "stage name");

    printf(fails ? "host checks: %d FAILED\n" : "host checks: ok\n", fails);
    return fails != 0;
}

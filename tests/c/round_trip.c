#include <stdlib.h>
#include "s2kit/FST_semi_memo.h"
#include "s2kit/cospml.h"
#include "s2kit/weights.h"

void round_trip(int bw, double* rdata, double* idata, double* rc, double* ic) {
    const int cutoff = bw;
    double* ws = malloc(sizeof(double) * (10 * bw * bw + 24 * bw));  /* not used by this library */
    double* tsp = malloc(sizeof(double) * Reduced_Naive_TableSize(bw, cutoff)
                        + sizeof(double) * Reduced_SpharmonicTableSize(bw, cutoff));
    double* ttsp = malloc(sizeof(double) * Reduced_Naive_TableSize(bw, cutoff)
                         + sizeof(double) * Reduced_SpharmonicTableSize(bw, cutoff));
    double* w = malloc(sizeof(double) * 4 * bw);
    double** t = SemiNaive_Naive_Pml_Table(bw, cutoff, tsp, ws);
    double** tt = Transpose_SemiNaive_Naive_Pml_Table(t, bw, cutoff, ttsp, ws);
    GenerateWeightsForDLT(bw, w);
    InvFSTSemiMemo(rc, ic, rdata, idata, bw, tt, ws, COMPLEX, cutoff, NULL, NULL);
    FSTSemiMemo(rdata, idata, rc, ic, bw, t, ws, COMPLEX, cutoff, NULL, NULL, w);
    free(t); free(tt);                   /* malloc'd pointer arrays, as in the reference */
    free(tsp); free(ttsp); free(w); free(ws);
}

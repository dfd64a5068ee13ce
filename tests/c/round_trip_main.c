/* Runs round_trip() (the C example of the integration guide, round_trip.c,
 * compiled as it is written there) on coefficients read from a file.
 *
 *   round_trip <bw> <in.bin> <out.bin>
 *
 * in.bin:  rcoeffs, icoeffs (bw*bw doubles each)
 * out.bin: rdata, idata (4*bw*bw doubles each: the InvFSTSemiMemo grid),
 *          then rcoeffs, icoeffs after FSTSemiMemo of that grid.
 */
#include <stdio.h>
#include <stdlib.h>

void round_trip(int bw, double* rdata, double* idata, double* rc, double* ic);

static int io(FILE* f, double* p, size_t n, int rd) {
    return (rd ? fread(p, sizeof(double), n, f) : fwrite(p, sizeof(double), n, f)) == n;
}

int main(int argc, char** argv) {
    if (argc != 4) {
        fprintf(stderr, "usage: %s bw in.bin out.bin\n", argv[0]);
        return 2;
    }
    const int bw = atoi(argv[1]);
    if (bw < 2 || bw > 1024) {
        fprintf(stderr, "bw must be in [2, 1024]\n");
        return 2;
    }
    const size_t nc = (size_t)bw * bw, ng = 4 * nc;
    double* rc = malloc(sizeof(double) * nc);
    double* ic = malloc(sizeof(double) * nc);
    double* rd = malloc(sizeof(double) * ng);
    double* id = malloc(sizeof(double) * ng);
    if (!rc || !ic || !rd || !id) return 3;
    FILE* f = fopen(argv[2], "rb");
    if (!f || !io(f, rc, nc, 1) || !io(f, ic, nc, 1)) {
        fprintf(stderr, "cannot read %s\n", argv[2]);
        return 3;
    }
    fclose(f);
    round_trip(bw, rd, id, rc, ic);
    f = fopen(argv[3], "wb");
    if (!f || !io(f, rd, ng, 0) || !io(f, id, ng, 0) || !io(f, rc, nc, 0) || !io(f, ic, nc, 0)) {
        fprintf(stderr, "cannot write %s\n", argv[3]);
        return 3;
    }
    fclose(f);
    free(rc);
    free(ic);
    free(rd);
    free(id);
    return 0;
}

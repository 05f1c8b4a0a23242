/*
 * A C caller of libsptol_amd.so written the way the reference's drivers are
 * (examples/mttkrp.c, mttkrp_hicoo.c, cpd.c): getopt options parsed with
 * the PARTI_SCN_* conversions, every call under sptAssert, status reports,
 * CPU timers.  It loads a .tns tensor, sorts it, runs sptMTTKRP and
 * sptMTTKRPHiCOO on every mode and sptCpdAls, and writes each input and
 * result with full FP32 precision into a directory for the test that runs
 * it (tests/test_gpu_drivers.py) to compare with the oracle.
 *
 *   sptol_driver -i X.tns -o OUTDIR [-r RANK] [-n CPD_ITERS] [-s SEED]
 *                [-b SB_BITS] [-k SK_BITS]
 */
#define _GNU_SOURCE
#include <getopt.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <ParTI.h>
#include <sptol_hip.h>

static FILE *open_out(const char *dir, const char *name)
{
    char *path = NULL;
    sptAssert(asprintf(&path, "%s/%s", dir, name) > 0);
    FILE *fp = fopen(path, "w");
    sptAssert(fp != NULL);
    free(path);
    return fp;
}

static void write_matrix(FILE *fp, const sptMatrix *m, sptIndex nrows)
{
    for (sptIndex i = 0; i < nrows; ++i) {
        for (sptIndex j = 0; j < m->ncols; ++j)
            fprintf(fp, "%.9e%c", m->values[(size_t)i * m->stride + j], j + 1 < m->ncols ? ' ' : '\n');
    }
}

int main(int argc, char **argv)
{
    const char *in = NULL, *outdir = NULL;
    sptIndex R = 16, niters = 5;
    sptElementIndex sb_bits = 7, sk_bits = 10;
    unsigned long long seed = 2024;
    int c;
    while ((c = getopt(argc, argv, "i:o:r:n:s:b:k:")) != -1) {
        switch (c) {
        case 'i': in = optarg; break;
        case 'o': outdir = optarg; break;
        case 'r': sscanf(optarg, "%" PARTI_SCN_INDEX, &R); break;
        case 'n': sscanf(optarg, "%" PARTI_SCN_INDEX, &niters); break;
        case 's': sscanf(optarg, "%llu", &seed); break;
        case 'b': sscanf(optarg, "%" PARTI_SCN_ELEMENT_INDEX, &sb_bits); break;
        case 'k': sscanf(optarg, "%" PARTI_SCN_ELEMENT_INDEX, &sk_bits); break;
        default:
            fprintf(stderr, "usage: %s -i X.tns -o OUTDIR [-r R] [-n ITERS] [-s SEED] [-b SB] [-k SK]\n", argv[0]);
            return 1;
        }
    }
    sptAssert(in != NULL && outdir != NULL);

    sptSparseTensor X;
    FILE *fi = fopen(in, "r");
    sptAssert(fi != NULL);
    sptAssert(sptLoadSparseTensor(&X, 1, fi) == 0);
    fclose(fi);
    sptSparseTensorSortIndex(&X, 1, 1);
    sptSparseTensorStatus(&X, stdout);
    const sptIndex N = X.nmodes;

    /* the tensor in the order every COO call below sees it */
    FILE *ft = open_out(outdir, "tensor.txt");
    for (sptNnzIndex z = 0; z < X.nnz; ++z) {
        for (sptIndex m = 0; m < N; ++m) fprintf(ft, "%" PARTI_PRI_INDEX " ", X.inds[m].data[z]);
        fprintf(ft, "%.9e\n", X.values.data[z]);
    }
    fclose(ft);

    /* factors from the library's seeded generator */
    sptolSetSeed(seed);
    sptMatrix **U = (sptMatrix **)malloc((N + 1) * sizeof(sptMatrix *));
    sptIndex max_dim = 0;
    FILE *ff = open_out(outdir, "factors.txt");
    for (sptIndex m = 0; m < N; ++m) {
        U[m] = (sptMatrix *)malloc(sizeof(sptMatrix));
        sptAssert(sptNewMatrix(U[m], X.ndims[m], R) == 0);
        sptAssert(sptRandomizeMatrix(U[m], X.ndims[m], R) == 0);
        write_matrix(ff, U[m], X.ndims[m]);
        if (X.ndims[m] > max_dim) max_dim = X.ndims[m];
    }
    fclose(ff);
    U[N] = (sptMatrix *)malloc(sizeof(sptMatrix));
    sptAssert(sptNewMatrix(U[N], max_dim, R) == 0);
    sptAssert(sptConstantMatrix(U[N], 0) == 0);
    sptIndex *mats_order = (sptIndex *)malloc(N * sizeof(sptIndex));

    for (sptIndex mode = 0; mode < N; ++mode) {
        for (sptIndex k = 0; k < N; ++k) mats_order[k] = (mode + k) % N;
        sptTimer timer;
        sptAssert(sptNewTimer(&timer, 0) == 0);
        sptAssert(sptStartTimer(timer) == 0);
        for (int it = 0; it < 3; ++it) sptAssert(sptMTTKRP(&X, U, mats_order, mode) == 0);
        sptAssert(sptStopTimer(timer) == 0);
        char *name = NULL;
        sptAssert(asprintf(&name, "GPU  SpTns MTTKRP MODE %" PARTI_PRI_INDEX, mode) > 0);
        sptPrintAverageElapsedTime(timer, 3, name);
        free(name);
        sptFreeTimer(timer);
        char *fname = NULL;
        sptAssert(asprintf(&fname, "coo_mode%" PARTI_PRI_INDEX ".txt", mode) > 0);
        FILE *fo = open_out(outdir, fname);
        write_matrix(fo, U[N], X.ndims[mode]);
        fclose(fo);
        free(fname);
    }

    /* HiCOO over a copy (the conversion permutes its input) */
    sptSparseTensor Y;
    sptAssert(sptCopySparseTensor(&Y, &X, 1) == 0);
    sptSparseTensorHiCOO H;
    sptNnzIndex max_nnzb = 0;
    sptAssert(sptSparseTensorToHiCOO(&H, &max_nnzb, &Y, sb_bits, sk_bits, 1) == 0);
    sptSparseTensorStatusHiCOO(&H, stdout);
    for (sptIndex mode = 0; mode < N; ++mode) {
        for (sptIndex k = 0; k < N; ++k) mats_order[k] = (mode + k) % N;
        sptAssert(sptMTTKRPHiCOO(&H, U, mats_order, mode) == 0);
        char *fname = NULL;
        sptAssert(asprintf(&fname, "hicoo_mode%" PARTI_PRI_INDEX ".txt", mode) > 0);
        FILE *fo = open_out(outdir, fname);
        write_matrix(fo, U[N], X.ndims[mode]);
        fclose(fo);
        free(fname);
    }
    sptFreeSparseTensorHiCOO(&H);
    sptFreeSparseTensor(&Y);

    /* CPD-ALS: the initial factors come from the generator reseeded with
     * seed + 1, all niters iterations (tol 0) */
    sptKruskalTensor kt;
    sptAssert(sptNewKruskalTensor(&kt, N, X.ndims, R) == 0);
    sptolSetSeed(seed + 1);
    sptAssert(sptCpdAls(&X, R, niters, 0.0, &kt) == 0);
    FILE *fc = open_out(outdir, "cpd.txt");
    fprintf(fc, "%.12e\n", kt.fit);
    for (sptIndex r = 0; r < R; ++r) fprintf(fc, "%.9e%c", kt.lambda[r], r + 1 < R ? ' ' : '\n');
    for (sptIndex m = 0; m < N; ++m) write_matrix(fc, kt.factors[m], X.ndims[m]);
    fclose(fc);
    printf("fit: %.9f\n", kt.fit);

    sptFreeKruskalTensor(&kt);
    for (sptIndex m = 0; m <= N; ++m) { sptFreeMatrix(U[m]); free(U[m]); }
    free(U);
    free(mats_order);
    sptFreeSparseTensor(&X);
    printf("driver done\n");
    return 0;
}

/* rgf_inverse_driver: runs the slice-inverse kernels through the C ABI
 * (rgf_dense_inverse) on an n x n complex matrix with a dominant diagonal,
 * the shape of one Dyson step of an n-channel slab.  A plain host program
 * (no Python), used as the target of hardware-counter runs.
 *
 * usage: rgf_inverse_driver [n=4096] [precision=1 (0 FP64, 1 FP32 kernels)] [reps=1]
 *        [3m=0 (1: FP32 kernels with three real products per complex product)]
 */
#include <stdio.h>
#include <stdlib.h>

#include "rgf.h"

int main(int argc, char** argv) {
    const int64_t n = argc > 1 ? atoll(argv[1]) : 4096;
    const int prec = argc > 2 ? atoi(argv[2]) : RGF_PRECISION_FP32_REFINED;
    const int reps = argc > 3 ? atoi(argv[3]) : 1;
    const int use3m = argc > 4 ? atoi(argv[4]) : 0;
    if (n <= 0 || reps <= 0) {
        fprintf(stderr, "usage: %s [n] [precision] [reps]\n", argv[0]);
        return 1;
    }
    rgf_handle h;
    if (rgf_create(&h, 0) != RGF_OK) {
        fprintf(stderr, "rgf_create failed: %s\n", rgf_last_error(NULL));
        return 1;
    }
    double* A = (double*)malloc(sizeof(double) * 2 * (size_t)n * (size_t)n);
    if (!A) return 1;
    srand(1);
    for (int r = 0; r < reps; ++r) {
        for (int64_t e = 0; e < n * n; ++e) {
            A[2 * e] = 0.05 * ((rand() % 2001) / 1000.0 - 1.0);
            A[2 * e + 1] = 0.05 * ((rand() % 2001) / 1000.0 - 1.0);
        }
        for (int64_t i = 0; i < n; ++i) A[2 * (i * n + i)] += 6.0;
        const rgf_status st = rgf_dense_inverse(h, use3m ? (prec | RGF_DENSE_3M) : prec, n, A, 0);
        printf("rep %d: %s\n", r, st == RGF_OK ? "ok" : rgf_last_error(h));
        fflush(stdout);
        if (st != RGF_OK) return 2;
    }
    rgf_destroy(h);
    free(A);
    return 0;
}

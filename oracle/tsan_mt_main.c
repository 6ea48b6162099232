/* ThreadSanitizer driver for the threaded CPU baseline (test
 * infrastructure only): oracle_time_spmv_csr_f64_mt on a random CSR with
 * several worker counts, each result compared with the serial
 * restatement.  Built and run by `make -C oracle tsan`; exits 1 on a
 * mismatch (ThreadSanitizer reports races itself, with exit code 66). */
#include <stdio.h>
#include <stdlib.h>

double oracle_time_spmv_csr_f64_mt(int nrows, const int *rowptr, const int *col, const double *val,
                                   const double *x, double *y, int workers, int warm, int reps);
void oracle_spmv_csr_f64(int N, const int *rowptr, const int *col, const double *val, const double *x, double *y,
                         int inner_max);

int main(void)
{
    const int N = 20000;
    int *rp = malloc(sizeof(int) * (N + 1));
    unsigned s = 12345u;
    rp[0] = 0;
    for (int i = 0; i < N; i++) {
        s = s * 1103515245u + 12345u;
        rp[i + 1] = rp[i] + (int)((s >> 16) % 40u) + 1;
    }
    const int nnz = rp[N];
    int *col = malloc(sizeof(int) * nnz);
    double *val = malloc(sizeof(double) * nnz), *x = malloc(sizeof(double) * N);
    double *y = malloc(sizeof(double) * N), *yr = calloc(N, sizeof(double));
    for (int k = 0; k < nnz; k++) {
        s = s * 1103515245u + 12345u;
        col[k] = (int)((s >> 8) % (unsigned)N);
        val[k] = (double)((s >> 4) % 1000u) / 500.0 - 1.0;
    }
    for (int i = 0; i < N; i++) x[i] = (double)(i % 17) / 8.0 - 1.0;
    oracle_spmv_csr_f64(N, rp, col, val, x, yr, 1);
    int bad = 0;
    const int workers[] = {1, 3, 8};
    for (int w = 0; w < 3; w++) {
        oracle_time_spmv_csr_f64_mt(N, rp, col, val, x, y, workers[w], 1, 3);
        for (int i = 0; i < N; i++)
            if (y[i] != yr[i]) bad = 1;
    }
    printf("threaded baseline: %s\n", bad ? "MISMATCH" : "equal to the serial restatement");
    free(rp); free(col); free(val); free(x); free(y); free(yr);
    return bad;
}

/*
 * A plain C11 caller of libptd through include/ptd.h, built by tests/test_c_abi.py with gcc (no C++, no HIP
 * headers) into a small shared library that the test calls: it runs the physical-time loop of a script
 * (ptd_solve) from H0 and returns the iteration counts, the errors and the final H, as a Julia `ccall`
 * binding or any other C host would.
 */
#include <stddef.h>
#include <stdio.h>

#include "ptd.h"

static int32_t report(const char* what, int32_t st) {
    fprintf(stderr, "abi_solve: %s failed (status %d): %s\n", what, (int)st, ptd_last_error());
    return st;
}

int32_t abi_solve(const ptd_params* params, const double* h0, int32_t nt, int64_t* iters, double* errs,
                  double* h_out) {
    const int64_t n = params->nx * params->ny * params->nz;
    ptd_ctx* ctx = NULL;
    int32_t st = ptd_ctx_create(params, 0, &ctx);
    if (st != PTD_OK) return report("ptd_ctx_create", st);
    if ((st = ptd_set_field(ctx, PTD_F_H, h0, n)) != PTD_OK) {
        report("ptd_set_field(H)", st);
    } else if ((st = ptd_set_field(ctx, PTD_F_HOLD, h0, n)) != PTD_OK) {
        report("ptd_set_field(Hold)", st);
    } else if ((st = ptd_solve(ctx, nt, iters, errs)) != PTD_OK) {
        report("ptd_solve", st);
    } else if ((st = ptd_get_field(ctx, PTD_F_H, h_out, n)) != PTD_OK) {
        report("ptd_get_field(H)", st);
    }
    const int32_t st2 = ptd_ctx_destroy(ctx);
    return st != PTD_OK ? st : st2;
}

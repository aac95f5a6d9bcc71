/*
 * MATLAB gateway `ph` over libopenph (gfx950).
 *
 * Same name, arity and outputs as the reference's gateway
 * (src/matlab/ph/ph.cu:11-92), so openph.m (src/matlab/ph/openph.m:3-21)
 * and every numerics/ caller keep working unchanged:
 *
 *   [low, ess, err_linf, err_lone, err_redu, err_ess, time_track, num_iters] =
 *       ph(algstr, int32 r[nnz x 1], int32 c[nnz x 1], int32 m,
 *          int32 col_width, int32 low_true[1 x m])
 *
 *   low         1 x m int32, 1-based, 0 = the reduced column is zero
 *   ess         1 x m int32, 1 = essential
 *   err_*       1 x (m+1) single, per-iteration trackers
 *   time_track  1 x (m+1) single, ms per iteration
 *   num_iters   1 x 1 int32 (negative status code on failure)
 *
 * Differences from the reference gateway, all on error paths:
 *   - wrong arity, a non-int32 array or a non-string algstr raises a MATLAB
 *     error instead of printing and returning unassigned outputs
 *     (src/matlab/ph/ph.cu:13-14);
 *   - a failed reduction keeps the outputs as initialised below, stores the
 *     negative status in num_iters and issues a warning naming the cause.
 *
 * Build (MATLAB R2018a or later, next to libopenph.so built by
 * openph_amd/csrc/Makefile):
 *   mex -R2018a -I<repo>/include ph_mex.c -L<repo>/openph_amd -lopenph \
 *       LDFLAGS='$LDFLAGS -Wl,-rpath,<repo>/openph_amd' -output ph
 */
#include <string.h>

#include "mex.h"
#include "openph.h"

static const int *int32_in(const mxArray *a, const char *what) {
    if (!mxIsInt32(a) || mxIsComplex(a))
        mexErrMsgIdAndTxt("openph:ph:type", "%s must be a real int32 array", what);
    return (const int *)mxGetData(a);
}

static int int_scalar(const mxArray *a, const char *what) {
    if (!mxIsNumeric(a) || mxGetNumberOfElements(a) != 1)
        mexErrMsgIdAndTxt("openph:ph:type", "%s must be a numeric scalar", what);
    return (int)mxGetScalar(a);
}

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]) {
    char algstr[64];
    int i;
    if (nrhs != 6 || nlhs > 8)
        mexErrMsgIdAndTxt("openph:ph:arity",
                          "usage: [low, ess, err_linf, err_lone, err_redu, err_ess, "
                          "time_track, num_iters] = ph(algstr, r, c, m, col_width, low_true)");
    if (!mxIsChar(prhs[0]) || mxGetString(prhs[0], algstr, sizeof(algstr)) != 0)
        mexErrMsgIdAndTxt("openph:ph:algstr", "algstr must be a short character vector");

    const int *rows = int32_in(prhs[1], "r");
    const int *cols = int32_in(prhs[2], "c");
    const int m = int_scalar(prhs[3], "m");
    const int col_width = int_scalar(prhs[4], "col_width");
    const int *low_true = int32_in(prhs[5], "low_true");
    /* nnz is the row count of r (src/matlab/ph/ph.cu:40) */
    const int nnz = (int)mxGetM(prhs[1]);
    if (m < 0) mexErrMsgIdAndTxt("openph:ph:m", "m must be non-negative");
    if ((int)mxGetNumberOfElements(prhs[2]) != nnz)
        mexErrMsgIdAndTxt("openph:ph:nnz", "r and c must have the same length");
    if ((int)mxGetNumberOfElements(prhs[5]) < m)
        mexErrMsgIdAndTxt("openph:ph:low_true", "low_true must have m entries");

    mxArray *out[8];
    out[0] = mxCreateNumericMatrix(1, m, mxINT32_CLASS, mxREAL);
    out[1] = mxCreateNumericMatrix(1, m, mxINT32_CLASS, mxREAL);
    for (i = 2; i < 7; i++) out[i] = mxCreateNumericMatrix(1, m + 1, mxSINGLE_CLASS, mxREAL);
    out[7] = mxCreateNumericMatrix(1, 1, mxINT32_CLASS, mxREAL);

    int *low = (int *)mxGetData(out[0]);
    int *ess = (int *)mxGetData(out[1]);
    float *err_linf = (float *)mxGetData(out[2]);
    float *err_lone = (float *)mxGetData(out[3]);
    float *err_redu = (float *)mxGetData(out[4]);
    float *err_ess = (float *)mxGetData(out[5]);
    float *time_track = (float *)mxGetData(out[6]);
    int *num_iters = (int *)mxGetData(out[7]);

    /* the reference gateway's initial values (src/matlab/ph/ph.cu:72-78) */
    for (i = 0; i < m; i++) low[i] = -1;
    for (i = 0; i < m; i++) ess[i] = -1;
    for (i = 0; i < m + 1; i++) {
        err_linf[i] = -1.0f;
        err_lone[i] = -1.0f;
        err_redu[i] = -1.0f;
        err_ess[i] = -1.0f;
        time_track[i] = 0.0f;
    }

    /* the reference fixes the device at 0 (src/matlab/ph/ph.cu:26) */
    openph(algstr, (int *)rows, (int *)cols, m, col_width, (int *)low_true, nnz, low, ess,
           err_linf, err_lone, err_redu, err_ess, time_track, num_iters, 0);
    if (num_iters[0] < 0)
        mexWarnMsgIdAndTxt("openph:ph:status", "openph failed: %s",
                           openph_strerror(num_iters[0]));

    for (i = 0; i < 8 && i < (nlhs > 0 ? nlhs : 1); i++) plhs[i] = out[i];
    for (; i < 8; i++) mxDestroyArray(out[i]);
}

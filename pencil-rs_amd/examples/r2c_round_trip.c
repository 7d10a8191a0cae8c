/*
 * The C ABI on its own: a 3D R2C forward + normalized C2R inverse of a
 * 64 x 48 x 40 f32 field through include/pencil_fft_hip.h, with device
 * buffers from the HIP runtime (no Python, no torch), as a pencil-rs backend
 * would drive it (R2cPlan::forward / inverse, distributed/r2c.rs).
 * Prints one line: "r2c_round_trip rel_l2=<e> spectrum_dc=<re>" and exits 0
 * when the round trip is within 2e-6.
 */
#include <hip/hip_runtime_api.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "pencil_fft_hip.h"

#define CHECK_HIP(x)                                                           \
    do {                                                                       \
        hipError_t e_ = (x);                                                   \
        if (e_ != hipSuccess) {                                                \
            fprintf(stderr, "HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__); \
            return 2;                                                          \
        }                                                                      \
    } while (0)

static int check(pfh_status st, const char *what)
{
    if (st == PFH_OK) return 0;
    pfh_error_info info;
    pfh_last_error(&info);
    fprintf(stderr, "%s: %s (%s)\n", what, pfh_status_string(st), info.message);
    return 1;
}

int main(void)
{
    const size_t n0 = 64, n1 = 48, n2 = 40, n = n0 * n1 * n2;
    pfh_desc d = {0};
    d.kind = PFH_R2C;
    d.precision = PFH_F32;
    d.ndim = 3;
    d.shape[0] = n0;
    d.shape[1] = n1;
    d.shape[2] = n2;
    d.permute_dims = 1;
    pfh_plan plan = NULL;
    if (check(pfh_plan_create(&d, &plan), "plan")) return 1;

    int perm[PFH_MAX_DIMS];
    size_t shape[PFH_MAX_DIMS], out_elems = 0, out_bytes = 0, ws_bytes = 0;
    if (check(pfh_plan_layout(plan, 1, perm, shape, &out_elems, &out_bytes), "layout")) return 1;
    if (check(pfh_plan_workspace_bytes(plan, PFH_INVERSE, &ws_bytes), "workspace")) return 1;

    float *hx = (float *)malloc(n * sizeof(float)), *hb = (float *)malloc(n * sizeof(float));
    unsigned s = 12345u;
    for (size_t i = 0; i < n; ++i) {
        s = s * 1664525u + 1013904223u;
        hx[i] = (float)((s >> 8) * (1.0 / 16777216.0) - 0.5);
    }
    void *x = NULL, *spec = NULL, *back = NULL, *ws = NULL;
    CHECK_HIP(hipMalloc(&x, n * sizeof(float)));
    CHECK_HIP(hipMalloc(&spec, out_elems * out_bytes));
    CHECK_HIP(hipMalloc(&back, n * sizeof(float)));
    CHECK_HIP(hipMalloc(&ws, ws_bytes));
    CHECK_HIP(hipMemset(ws, 0, ws_bytes)); /* a new workspace starts zeroed */
    CHECK_HIP(hipMemcpy(x, hx, n * sizeof(float), hipMemcpyHostToDevice));

    if (check(pfh_execute(plan, PFH_FORWARD, x, n, spec, out_elems, ws, ws_bytes, NULL, 0), "forward")) return 1;
    if (check(pfh_execute(plan, PFH_INVERSE, spec, out_elems, back, n, ws, ws_bytes, NULL, 0), "inverse")) return 1;
    CHECK_HIP(hipDeviceSynchronize());
    CHECK_HIP(hipMemcpy(hb, back, n * sizeof(float), hipMemcpyDeviceToHost));
    float dc[2] = {0, 0};
    CHECK_HIP(hipMemcpy(dc, spec, sizeof(dc), hipMemcpyDeviceToHost));

    double num = 0, den = 0, sum = 0;
    for (size_t i = 0; i < n; ++i) {
        num += (double)(hb[i] - hx[i]) * (hb[i] - hx[i]);
        den += (double)hx[i] * hx[i];
        sum += hx[i];
    }
    const double err = sqrt(num / den);
    printf("r2c_round_trip rel_l2=%.3e spectrum_dc=%.6e sum=%.6e\n", err, dc[0], sum);
    pfh_plan_destroy(plan);
    hipFree(x);
    hipFree(spec);
    hipFree(back);
    hipFree(ws);
    free(hx);
    free(hb);
    /* the DC bin of the forward transform is the sum of the field */
    return (err <= 2e-6 && fabs(dc[0] - sum) <= 1e-4 * (fabs(sum) + 1.0)) ? 0 : 1;
}

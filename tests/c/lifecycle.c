/* lifecycle.c -- a C host driving libsle through include/sle_api.h only, the
 * way the reference's examples/example_usage.cpp drives its library: create,
 * load a network, push measurements, solve, hot-start, read state.
 *
 * The IEEE 14-bus loader sets every measurement value to 0; the telemetry of
 * the first solve is read from the file named by argv[1] (one value per
 * measurement, insertion order).
 *
 * Prints "nogpu" when sle_Initialize reports no device (host-only checks still
 * run); otherwise one "V <i> <value>" and "T <i> <value>" line per bus after a
 * PRECISION solve and a REALTIME re-solve with perturbed telemetry, then
 * "status <precision> <realtime> <iterations>".  Exit code 0 on success. */
#include <stdio.h>
#include <stdlib.h>

#include "sle_api.h"

#define CHECK(x)                                                         \
    do {                                                                 \
        SLE_StatusCode rc_ = (x);                                        \
        if (rc_ != SLE_OK) {                                             \
            fprintf(stderr, "%s -> %d (%s)\n", #x, (int)rc_,             \
                    sle_GetLastError(h) ? sle_GetLastError(h) : "");     \
            return 1;                                                    \
        }                                                                \
    } while (0)

int main(int argc, char** argv) {
    SLE_Handle h = NULL;
    if (sle_Create(&h) != SLE_OK) return 1;
    CHECK(sle_LoadIEEE14Bus(h));
    const int nb = sle_GetBusCount(h), nm = sle_GetMeasurementCount(h);
    if (nb != 14 || nm <= 0) return 2;
    if (sle_Initialize(h) != SLE_OK) {
        printf("nogpu\n");
        sle_Destroy(h);
        return 0;
    }
    CHECK(sle_UploadModel(h));
    SLE_Real* z = (SLE_Real*)malloc(sizeof(SLE_Real) * (size_t)nm);
    if (argc < 2) return 4;
    FILE* fp = fopen(argv[1], "r");
    if (!fp) return 4;
    for (int i = 0; i < nm; ++i) {
        double val;
        if (fscanf(fp, "%lf", &val) != 1) return 4;
        z[i] = (SLE_Real)val;
    }
    fclose(fp);
    CHECK(sle_UpdateTelemetry(h, z, nm));
    SLE_Result r1, r2;
    CHECK(sle_Solve(h, SLE_MODE_PRECISION, &r1));
    /* hot start: perturb the telemetry slightly and re-solve in real time */
    SLE_MeasurementDetails d;
    for (int i = 0; i < nm; ++i) {
        char id[64];
        const char* p = sle_GetMeasurementIdByIndex(h, i);  /* valid until the next call */
        if (!p) return 3;
        snprintf(id, sizeof(id), "%s", p);
        if (sle_GetMeasurementDetails(h, id, &d) != SLE_OK) return 3;
        z[i] = d.current_value * (SLE_Real)(1.0 + 1e-4 * ((i % 7) - 3));
    }
    CHECK(sle_UpdateTelemetry(h, z, nm));
    CHECK(sle_Solve(h, SLE_MODE_REALTIME, &r2));
    SLE_Real v[64], t[64];
    CHECK(sle_GetVoltageMagnitudes(h, v, nb));
    CHECK(sle_GetVoltageAngles(h, t, nb));
    for (int i = 0; i < nb; ++i) printf("V %d %.9g\nT %d %.9g\n", i, (double)v[i], i, (double)t[i]);
    printf("status %d %d %d\n", (int)r1.status, (int)r2.status, (int)r2.iterations);
    free(z);
    sle_Destroy(h);
    return 0;
}

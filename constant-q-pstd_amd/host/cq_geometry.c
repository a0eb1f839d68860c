/*
 * cq_geometry.c -- `ge`, the acquisition-geometry generator.
 *
 * Replaces geometry/geometry2d.c:7-152 of the reference: reads the same
 * parameters.txt (shot count, first shot x, shot interval, shot and receiver
 * depths, maximum offset, trace interval, dx, dz, nx, nz) and writes the same
 * files into the current directory: tracenum.bin (int32[shots]), shotxp.bin,
 * shotzp.bin (float32[shots]) and gxp<i>.bin, gzp<i>.bin (float32[nr_i]).
 * Receivers of shot i cover [max(0, x_i - offset), min(xmax, x_i + offset)]
 * every trace interval (double-sided spread), with the reference's float
 * arithmetic.  Errors return a non-zero exit status instead of exit(0).
 *
 * usage: ge [parameters.txt]
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static int label_value(FILE *fp, const char *fmt, void *v)
{
    char buf[256];
    if (fscanf(fp, " ") < 0 || !fgets(buf, sizeof buf, fp)) return -1;
    return fscanf(fp, fmt, v) == 1 ? 0 : -1;
}

static int write_bin(const char *name, const void *p, size_t sz, size_t n)
{
    FILE *f = fopen(name, "wb");
    if (!f) return -1;
    size_t w = fwrite(p, sz, n, f);
    fclose(f);
    return w == n ? 0 : -1;
}

int main(int argc, char **argv)
{
    const char *par = argc > 1 ? argv[1] : "parameters.txt";
    FILE *fp = fopen(par, "r");
    if (!fp) {
        fprintf(stderr, "Cannot open %s!\n", par);
        return 1;
    }
    int shotnum, nx, nz;
    float sx0, sxdx, sdep, rdep, offset, trdx, dx, dz;
    int e = 0;
    e |= label_value(fp, " %d", &shotnum);
    e |= label_value(fp, " %f", &sx0);
    e |= label_value(fp, " %f", &sxdx);
    e |= label_value(fp, " %f", &sdep);
    e |= label_value(fp, " %f", &rdep);
    e |= label_value(fp, " %f", &offset);
    e |= label_value(fp, " %f", &trdx);
    e |= label_value(fp, " %f", &dx);
    e |= label_value(fp, " %f", &dz);
    e |= label_value(fp, " %d", &nx);
    e |= label_value(fp, " %d", &nz);
    fclose(fp);
    if (e || shotnum < 1) {
        fprintf(stderr, "malformed %s\n", par);
        return 1;
    }
    int *tracenum = (int *)malloc(sizeof(int) * shotnum);
    float *sxp = (float *)malloc(sizeof(float) * shotnum), *szp = (float *)malloc(sizeof(float) * shotnum);
    const float xmax = (nx - 1) * dx, zmax = (nz - 1) * dz;
    for (int i = 0; i < shotnum; i++) {
        sxp[i] = sx0 + i * sxdx;
        szp[i] = sdep;
        if (sxp[i] < 0.0 || sxp[i] > xmax || szp[i] < 0.0 || szp[i] > zmax) {
            fprintf(stderr, "Error! shot position exceeds model scope!!!\n");
            return 1;
        }
        const float xlp = (sxp[i] - offset > 0.0) ? sxp[i] - offset : 0.0;
        const float xrp = (sxp[i] + offset < xmax) ? sxp[i] + offset : xmax;
        const int nr = (int)((xrp - xlp) / trdx) + 1;
        float *rx = (float *)malloc(sizeof(float) * nr), *rz = (float *)malloc(sizeof(float) * nr);
        for (int j = 0; j < nr; j++) {
            rx[j] = xlp + j * trdx;
            rz[j] = rdep;
            if (rx[j] < 0.0 || rx[j] > xmax || rz[j] < 0.0 || rz[j] > zmax) {
                fprintf(stderr, "Error! Receiver position exceeds model scope!!!\n");
                return 1;
            }
        }
        char name[64];
        snprintf(name, sizeof name, "gxp%d.bin", i);
        if (write_bin(name, rx, 4, nr)) return 1;
        snprintf(name, sizeof name, "gzp%d.bin", i);
        if (write_bin(name, rz, 4, nr)) return 1;
        tracenum[i] = nr;
        printf("id=%d nr=%d xlp=%f xrp=%f xdis=%f\n", i, nr, xlp, xrp, xrp - xlp);
        free(rx);
        free(rz);
    }
    if (write_bin("shotxp.bin", sxp, 4, shotnum) || write_bin("shotzp.bin", szp, 4, shotnum) ||
        write_bin("tracenum.bin", tracenum, 4, shotnum))
        return 1;
    free(tracenum);
    free(sxp);
    free(szp);
    return 0;
}

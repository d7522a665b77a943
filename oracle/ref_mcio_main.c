/* Test-only driver around the reference's own file routines (NMC/NMCCode/mcio.h),
 * compiled from the reference sources in place into oracle/_ref/ref_mcio
 * (see the `ref` target of oracle/Makefile).  It lets the tests compare
 * nmc_amd/mcio.py with the reference byte for byte:
 *
 *   ref_mcio read <name>        ReadRunParams (mcio.h:16-171) + ImportBinaryTissueFile
 *                               (mcio.h:173-189) + PrintRunParameters (mcio.h:191-296)
 *                               to stdout + SaveOpticalProperties (mcio.h:298-322);
 *                               writes the parsed structs to <name>_parsed.bin
 *                               (RunParams then TissueParams, raw bytes)
 *   ref_mcio savefile <blob>    SaveFile (mcio.h:327-389) into ./mcOUT0.dat; <blob> holds
 *                               16 doubles (S A E mua mus g n1 n2 radius waist xs ys zs
 *                               dr dz Nphotons), 3 int32 (mcflag NR NZ), then
 *                               J float[NR+1] and F float[(NR+1)*NR + NZ + 1]
 *
 * Never part of the product: only tests run it, and only where it was built. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "definitions.h"
#include "mcio.h"

static int cmd_read(const char* name)
{
    struct RunParams rp;
    struct TissueParams tp;
    memset(&rp, 0, sizeof rp);
    memset(&tp, 0, sizeof tp);
    const char* argv[2] = {"ref_mcio", name};
    ReadRunParams(argv, &rp, &tp);
    char* v = ImportBinaryTissueFile(&rp, &tp);
    PrintRunParameters(&rp, &tp, v);
    SaveOpticalProperties(&rp, &tp);
    free(v);
    char path[512];
    snprintf(path, sizeof path, "%s_parsed.bin", name);
    FILE* f = fopen(path, "wb");
    if (!f) return 2;
    fwrite(&rp, sizeof rp, 1, f);
    fwrite(&tp, sizeof tp, 1, f);
    fclose(f);
    fflush(stdout);
    return 0;
}

static int cmd_savefile(const char* blob)
{
    FILE* f = fopen(blob, "rb");
    if (!f) return 2;
    double d[16];
    int n[3];
    if (fread(d, sizeof(double), 16, f) != 16 || fread(n, sizeof(int), 3, f) != 3) { fclose(f); return 3; }
    long nr = n[1], nz = n[2];
    long nj = nr + 1, nf = (nr + 1) * nr + nz + 1;
    float* J = (float*)calloc((size_t)nj, sizeof(float));
    float* F = (float*)calloc((size_t)nf, sizeof(float));
    if (!J || !F || fread(J, sizeof(float), (size_t)nj, f) != (size_t)nj ||
        fread(F, sizeof(float), (size_t)nf, f) != (size_t)nf) {
        fclose(f);
        return 4;
    }
    fclose(f);
    SaveFile(0, J, F, d[0], d[1], d[2], d[3], d[4], d[5], d[6], d[7], (short)n[0], d[8], d[9],
             d[10], d[11], d[12], (short)nr, (short)nz, d[13], d[14], d[15]);
    free(J);
    free(F);
    return 0;
}

int main(int argc, char** argv)
{
    if (argc == 3 && strcmp(argv[1], "read") == 0) return cmd_read(argv[2]);
    if (argc == 3 && strcmp(argv[1], "savefile") == 0) return cmd_savefile(argv[2]);
    fprintf(stderr, "usage: ref_mcio read <name> | ref_mcio savefile <blob>\n");
    return 1;
}

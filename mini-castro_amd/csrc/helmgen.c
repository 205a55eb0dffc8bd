/*
 * Helmholtz electron free-energy table generator.
 *
 * mini-Castro's EOS reads a 541 x 201 table of the electron free energy and
 * its derivatives from `helm_table.dat` (Source/eos.F90:626-670).  That file is
 * not shipped with the reference (.MISSING_LARGE_BLOBS:1), so this library
 * generates a table in exactly the same layout:
 *
 *   section 1, j = 1..201 (temperature, outer), i = 1..541 (density, inner):
 *       f  fd  ft  fdd  ftt  fdt  fddt  fdtt  fddtt          (Source/eos.F90:644-649)
 *   section 2: dpdf dpdfd dpdft dpdfdt                       (Source/eos.F90:652-656)
 *   section 3: ef   efd   eft   efdt                         (Source/eos.F90:659-663)
 *   section 4: xf   xfd   xft   xfdt                         (Source/eos.F90:666-670)
 *
 * on the grid t(j) = 10**(3 + (j-1)*0.05) K, d(i) = 10**(-12 + (i-1)*0.05)
 * g/cm^3 of ye*rho (Source/eos.F90:60-69,626-632).
 *
 * Physics: a non-degenerate, non-relativistic ideal electron gas.  With
 * n = d*N_A electrons per cm^3 and thermal wavelength
 * lambda = h / sqrt(2 pi m_e k T), the free energy per gram divided by ye is
 *
 *     f(d, T) = N_A k T [ ln(n lambda^3 / 2) - 1 ]
 *             = A T (ln d - 1.5 ln T + C - 1),  A = k N_A,
 *       C = ln(N_A / 2) + 1.5 ln(h^2 / (2 pi m_e k)).
 *
 * The EOS forms p_e = d^2 f_d = n k T, e_e = ye (f - T f_t) = 1.5 ye N_A k T,
 * so every tabulated column has a closed form and the table is
 * thermodynamically consistent by construction.  All derivatives below are
 * the exact partial derivatives of f:
 *
 *   f   = A T L        f_d = A T/d        f_dd = -A T/d^2     f_t = A (L - 1.5)
 *   f_tt = -1.5 A/T
 *   f_dt = A/d         f_ddt = -A/d^2      f_dtt = 0           f_ddtt = 0
 *   dpdf = d/dd(d^2 f_d) = A T, dpdfd = 0, dpdft = A, dpdfdt = 0
 *   ef  = eta = L + 1 (the degeneracy parameter ln(n lambda^3/2)), efd = 1/d,
 *         eft = -1.5/T, efdt = 0
 *   xf  = N_A d,  xfd = N_A,  xft = 0,  xfdt = 0
 *
 * where L = ln d - 1.5 ln T + C - 1 = eta - 1.
 *
 * The generator is deterministic; text values are printed with 17
 * significant digits so that parsing the text gives back the same doubles
 * as the binary cache.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "helmgen.h"

/* Physical constants as used by the reference EOS (Source/eos.F90:129-133). */
static const double HG_H = 6.6260689633e-27;
static const double HG_AVO = 6.0221417930e23;
static const double HG_KERG = 1.380650424e-16;
/* Electron mass (CODATA 2006, the same vintage as the constants above). */
static const double HG_ME = 9.10938215e-28;

static double hg_tgrid(int j) { return pow(10.0, 3.0 + (double)j * 0.05); }
static double hg_dgrid(int i) { return pow(10.0, -12.0 + (double)i * 0.05); }

void mc_helm_fill(double *tab)
{
    const long nn = (long)MC_HELM_IMAX * MC_HELM_JMAX;
    const double A = HG_KERG * HG_AVO;
    const double pi = 3.14159265358979323846;
    const double C = log(HG_AVO / 2.0) +
                     1.5 * log(HG_H * HG_H / (2.0 * pi * HG_ME * HG_KERG));
    for (int j = 0; j < MC_HELM_JMAX; ++j) {
        const double T = hg_tgrid(j);
        for (int i = 0; i < MC_HELM_IMAX; ++i) {
            const double d = hg_dgrid(i);
            const double L = log(d) - 1.5 * log(T) + C - 1.0;
            const long o = (long)i + (long)MC_HELM_IMAX * j;
            double v[MC_HELM_NTAB];
            v[MC_HT_F] = A * T * L;
            v[MC_HT_FD] = A * T / d;
            v[MC_HT_FT] = A * (L - 1.5);
            v[MC_HT_FDD] = -A * T / (d * d);
            v[MC_HT_FTT] = -1.5 * A / T;
            v[MC_HT_FDT] = A / d;
            v[MC_HT_FDDT] = -A / (d * d);
            v[MC_HT_FDTT] = 0.0;
            v[MC_HT_FDDTT] = 0.0;
            v[MC_HT_DPDF] = A * T;
            v[MC_HT_DPDFD] = 0.0;
            v[MC_HT_DPDFT] = A;
            v[MC_HT_DPDFDT] = 0.0;
            v[MC_HT_EF] = L + 1.0;
            v[MC_HT_EFD] = 1.0 / d;
            v[MC_HT_EFT] = -1.5 / T;
            v[MC_HT_EFDT] = 0.0;
            v[MC_HT_XF] = HG_AVO * d;
            v[MC_HT_XFD] = HG_AVO;
            v[MC_HT_XFT] = 0.0;
            v[MC_HT_XFDT] = 0.0;
            for (int c = 0; c < MC_HELM_NTAB; ++c) tab[c * nn + o] = v[c];
        }
    }
}

/* Column order of the four text sections (Source/eos.F90:646-668). */
static const int hg_sec1[9] = {MC_HT_F, MC_HT_FD, MC_HT_FT, MC_HT_FDD, MC_HT_FTT,
                               MC_HT_FDT, MC_HT_FDDT, MC_HT_FDTT, MC_HT_FDDTT};
static const int hg_sec2[4] = {MC_HT_DPDF, MC_HT_DPDFD, MC_HT_DPDFT, MC_HT_DPDFDT};
static const int hg_sec3[4] = {MC_HT_EF, MC_HT_EFD, MC_HT_EFT, MC_HT_EFDT};
static const int hg_sec4[4] = {MC_HT_XF, MC_HT_XFD, MC_HT_XFT, MC_HT_XFDT};

int mc_helm_write_text(const double *tab, const char *path)
{
    FILE *fp = fopen(path, "w");
    if (!fp) return -1;
    const long nn = (long)MC_HELM_IMAX * MC_HELM_JMAX;
    const int *secs[4] = {hg_sec1, hg_sec2, hg_sec3, hg_sec4};
    const int ncol[4] = {9, 4, 4, 4};
    for (int s = 0; s < 4; ++s)
        for (long o = 0; o < nn; ++o) {
            for (int c = 0; c < ncol[s]; ++c)
                fprintf(fp, c ? " %.16e" : "%.16e", tab[secs[s][c] * nn + o]);
            fputc('\n', fp);
        }
    return fclose(fp) == 0 ? 0 : -1;
}

int mc_helm_read_text(double *tab, const char *path)
{
    FILE *fp = fopen(path, "r");
    if (!fp) return -1;
    const long nn = (long)MC_HELM_IMAX * MC_HELM_JMAX;
    const int *secs[4] = {hg_sec1, hg_sec2, hg_sec3, hg_sec4};
    const int ncol[4] = {9, 4, 4, 4};
    int rc = 0;
    for (int s = 0; s < 4 && rc == 0; ++s)
        for (long o = 0; o < nn && rc == 0; ++o)
            for (int c = 0; c < ncol[s]; ++c)
                if (fscanf(fp, "%lf", &tab[secs[s][c] * nn + o]) != 1) { rc = -2; break; }
    fclose(fp);
    return rc;
}

static const char hg_magic[8] = {'M', 'C', 'H', 'E', 'L', 'M', '0', '1'};

int mc_helm_write_bin(const double *tab, const char *path)
{
    FILE *fp = fopen(path, "wb");
    if (!fp) return -1;
    int dims[2] = {MC_HELM_IMAX, MC_HELM_JMAX};
    const size_t n = (size_t)MC_HELM_NTAB * MC_HELM_IMAX * MC_HELM_JMAX;
    int ok = fwrite(hg_magic, 1, 8, fp) == 8 && fwrite(dims, sizeof(int), 2, fp) == 2 &&
             fwrite(tab, sizeof(double), n, fp) == n;
    return (fclose(fp) == 0 && ok) ? 0 : -1;
}

int mc_helm_read_bin(double *tab, const char *path)
{
    FILE *fp = fopen(path, "rb");
    if (!fp) return -1;
    char magic[8];
    int dims[2];
    const size_t n = (size_t)MC_HELM_NTAB * MC_HELM_IMAX * MC_HELM_JMAX;
    int ok = fread(magic, 1, 8, fp) == 8 && memcmp(magic, hg_magic, 8) == 0 &&
             fread(dims, sizeof(int), 2, fp) == 2 && dims[0] == MC_HELM_IMAX &&
             dims[1] == MC_HELM_JMAX && fread(tab, sizeof(double), n, fp) == n;
    fclose(fp);
    return ok ? 0 : -2;
}

int mc_helm_read(double *tab, const char *path)
{
    size_t L = strlen(path);
    if (L > 4 && strcmp(path + L - 4, ".bin") == 0) return mc_helm_read_bin(tab, path);
    return mc_helm_read_text(tab, path);
}

#ifdef MC_HELMGEN_MAIN
int main(int argc, char **argv)
{
    if (argc < 2) {
        fprintf(stderr, "usage: %s helm_table.dat [helm_table.bin]\n", argv[0]);
        return 2;
    }
    const size_t n = (size_t)MC_HELM_NTAB * MC_HELM_IMAX * MC_HELM_JMAX;
    double *tab = (double *)malloc(n * sizeof(double));
    if (!tab) return 1;
    mc_helm_fill(tab);
    if (mc_helm_write_text(tab, argv[1]) != 0) { fprintf(stderr, "cannot write %s\n", argv[1]); return 1; }
    if (argc > 2 && mc_helm_write_bin(tab, argv[2]) != 0) { fprintf(stderr, "cannot write %s\n", argv[2]); return 1; }
    free(tab);
    return 0;
}
#endif

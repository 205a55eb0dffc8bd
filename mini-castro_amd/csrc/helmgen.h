/* Helmholtz table generator and readers (layout of Source/eos.F90:626-670). */
#ifndef MC_HELMGEN_H
#define MC_HELMGEN_H

#ifdef __cplusplus
extern "C" {
#endif

#define MC_HELM_IMAX 541
#define MC_HELM_JMAX 201
#define MC_HELM_NTAB 21

/* Column indices, in the order the reference declares its tables
   (Source/eos.F90:82-94). Each column is imax*jmax doubles, i fastest. */
enum {
    MC_HT_F = 0, MC_HT_FD, MC_HT_FT, MC_HT_FDD, MC_HT_FTT, MC_HT_FDT, MC_HT_FDDT,
    MC_HT_FDTT, MC_HT_FDDTT,
    MC_HT_DPDF, MC_HT_DPDFD, MC_HT_DPDFT, MC_HT_DPDFDT,
    MC_HT_EF, MC_HT_EFD, MC_HT_EFT, MC_HT_EFDT,
    MC_HT_XF, MC_HT_XFD, MC_HT_XFT, MC_HT_XFDT
};

/* Fill tab[21 * 541 * 201] with the generated ideal-electron-gas table. */
void mc_helm_fill(double *tab);
/* Text in the reference's helm_table.dat layout; binary cache "MCHELM01". */
int mc_helm_write_text(const double *tab, const char *path);
int mc_helm_read_text(double *tab, const char *path);
int mc_helm_write_bin(const double *tab, const char *path);
int mc_helm_read_bin(double *tab, const char *path);
/* Reads binary if path ends in ".bin", text otherwise. 0 on success. */
int mc_helm_read(double *tab, const char *path);

#ifdef __cplusplus
}
#endif
#endif
